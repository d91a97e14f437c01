// Cross-lane moves inside one wave64 on gfx950, as VALU operations: DPP row shifts inside a
// row of 16 lanes, and the lane swaps v_permlane16_swap (rows 2k <-> 2k+1) and
// v_permlane32_swap (lanes l <-> l + 32). A 4x4x4 block maps cell (x, y, z) to lane
// 16 x + 4 y + z, so its z and y neighbours are DPP shifts by 1 and 4 and its x neighbours
// lie 16 lanes away. Unlike ds_bpermute these do not go through the LDS crossbar.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace natray {

// Value of lane l + N (1 <= N <= 15) where that lane is in the same row of 16; lanes whose
// source falls outside their row get their own value back.
template <int N>
__device__ __forceinline__ uint32_t lane_up(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, 0x100 + N, 0xf, 0xf, false);
}
// Value of lane l - N, likewise.
template <int N>
__device__ __forceinline__ uint32_t lane_down(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, 0x110 + N, 0xf, 0xf, false);
}
// Value of lane l ^ 16.
__device__ __forceinline__ uint32_t lane_xor16(uint32_t x) {
  const auto r = __builtin_amdgcn_permlane16_swap(x, x, false, false);
  return (__lane_id() & 16u) ? r[0] : r[1];
}
// Value of lane l ^ 32.
__device__ __forceinline__ uint32_t lane_xor32(uint32_t x) {
  const auto r = __builtin_amdgcn_permlane32_swap(x, x, false, false);
  return (__lane_id() & 32u) ? r[0] : r[1];
}

// The same moves on 32- and 64-bit floats.
template <int N>
__device__ __forceinline__ float lane_up(float x) { return __uint_as_float(lane_up<N>(__float_as_uint(x))); }
template <int N>
__device__ __forceinline__ float lane_down(float x) { return __uint_as_float(lane_down<N>(__float_as_uint(x))); }
__device__ __forceinline__ float lane_xor16(float x) { return __uint_as_float(lane_xor16(__float_as_uint(x))); }
__device__ __forceinline__ float lane_xor32(float x) { return __uint_as_float(lane_xor32(__float_as_uint(x))); }

template <int N>
__device__ __forceinline__ double lane_up(double x) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  const uint32_t lo = lane_up<N>((uint32_t)b), hi = lane_up<N>((uint32_t)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ double lane_xor16(double x) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  const uint32_t lo = lane_xor16((uint32_t)b), hi = lane_xor16((uint32_t)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ double lane_xor32(double x) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  const uint32_t lo = lane_xor32((uint32_t)b), hi = lane_xor32((uint32_t)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Tree sum over the wave in the order v[l] += v[l + off] for off = 32, 16, ..., 1 (the
// block_sum tree of fluid.wgsl:414-422); lane 0 ends with the sum, other lanes hold partial
// or meaningless values.
template <class R>
__device__ __forceinline__ R lane_tree(R v) {
  v += lane_xor32(v);
  v += lane_xor16(v);
  v += lane_up<8>(v);
  v += lane_up<4>(v);
  v += lane_up<2>(v);
  v += lane_up<1>(v);
  return v;
}

// The six face neighbours of each lane's cell inside its 4x4x4 block. A neighbour outside the
// block gives a meaningless value; callers mask those.
struct Faces {
  float xm, xp, ym, yp, zm, zp;
};
__device__ __forceinline__ Faces block_faces(float v) {
  const uint32_t lx = __lane_id() >> 4;
  const float p16 = lane_xor16(v);    // the partner row: x + 1 for even x, x - 1 for odd x
  const float p48 = lane_xor32(p16);  // x + 1 for x = 1, x - 1 for x = 2
  Faces f;
  f.xm = (lx & 1u) ? p16 : p48;
  f.xp = (lx & 1u) ? p48 : p16;
  f.ym = lane_down<4>(v);
  f.yp = lane_up<4>(v);
  f.zm = lane_down<1>(v);
  f.zp = lane_up<1>(v);
  return f;
}

}  // namespace natray
