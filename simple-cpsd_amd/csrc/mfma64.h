/*
 * mfma64.h -- the FP64 MFMA building block shared by the GEMM kernels and the
 * Lagrange iteration: v_mfma_f64_4x4x4_4b_f64 and the rotated B operands that
 * let four of them cover a 16 x 16 output tile.
 *
 * Instruction layout (measured on gfx950): with A lane l = A[l%16][l/16] and
 * B lane l = B[l/16][l%16], one instruction forms the four diagonal 4x4
 * blocks of the 16x16 product A(16x4) B(4x16); output lane o = 16q + 4a + b
 * holds C[4a+q][4a+b].  Feeding B with its 4-column groups rotated by s
 * (column c taken from 4((c/4 + s) % 4) + c%4) gives C[4a+q][4((a+s)%4)+b],
 * so four rotations cover a 16 x 16 tile.
 */
#pragma once
#include <hip/hip_runtime.h>

__device__ __forceinline__ double mfma4(double a, double b, double c)
{
   return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0);
}

/* v taken from the lane L places along this lane's row of 16 (DPP row
 * rotate, two 32-bit moves); which way the row turns is read once per kernel
 * by rot_step() so that the epilogue indexing never depends on it */
template<int L>
__device__ __forceinline__ double row_ror(double v)
{
   const long long u = __double_as_longlong(v);
   const int lo = __builtin_amdgcn_mov_dpp((int)u, 0x120 + L, 0xf, 0xf, false);
   const int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), 0x120 + L, 0xf, 0xf, false);
   return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

/* 1 if row_ror<4> hands lane c the value of lane c+4, 3 if that of lane c-4;
 * the rotation fed by row_ror<4r> is then (rot_step() * r) % 4 */
__device__ __forceinline__ int rot_step(int lane)
{
   const int src = __builtin_amdgcn_mov_dpp(lane, 0x124, 0xf, 0xf, false);
   return ((src - lane) & 15) >> 2;
}

/* B operands for the four rotations of a 16-column block from the unrotated one */
__device__ __forceinline__ void rotations(double b0, double b[4])
{
   b[0] = b0;
   b[1] = row_ror<4>(b0);
   b[2] = row_ror<8>(b0);
   b[3] = row_ror<12>(b0);
}
