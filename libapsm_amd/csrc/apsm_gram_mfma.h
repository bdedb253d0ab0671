// Gram among the active samples of a train block on the matrix cores (FP32 inputs and
// accumulation). One wave computes a 32 x 32 block (rows i of block bi as the B operand, held
// in registers; rows j of block bj <= bi as the A operand, loaded from the stream buffer) of the
// two cross terms p = a_i.a_j and s = a_i.b_j with v_mfma_f32_32x32x2f32, and writes the lower
// triangle's kappa values; a workgroup takes four blocks.
#pragma once

#include "apsm_detect_mfma.h"

namespace apsm {

template <int NMAX>
__device__ __forceinline__ void gramBlockMfma( const float* __restrict__ Z, const float* __restrict__ Zn, float* __restrict__ Gact,
                                               const GramArgs& a, const int64_t blk )
{
    constexpr int K = 2 * NMAX;
    const int N = a.N, L = a.L;
    const int lane = threadIdx.x & 63;
    const int col = lane & 31, h = lane >> 5;
    // block (bi, bj), bj <= bi, from the linear lower-triangle index
    int64_t bi = (int64_t)( ( sqrt( 8.0 * (double)blk + 1.0 ) - 1.0 ) * 0.5 );
    while ( ( bi + 1 ) * ( bi + 2 ) / 2 <= blk )
        bi++;
    while ( bi * ( bi + 1 ) / 2 > blk )
        bi--;
    const int64_t bj = blk - bi * ( bi + 1 ) / 2;
    const int64_t i = bi * 32 + col; // this lane's row (B operand column)
    const int64_t jr = bj * 32 + col; // the row this lane supplies as A operand
    const bool vi = i < a.nA, vj = jr < a.nA;
    const float* xi = Z + ( a.c0 + i ) * L;
    const float* zj = Z + ( a.c0 + jr ) * L;

    // operands first (all loads in flight at once), then the MFMA chain
    float xv[ K / 2 ], zr[ K / 2 ], zim[ K / 2 ];
#pragma unroll
    for ( int t = 0; t < K / 2; t++ )
    {
        const int k = 2 * t + h;
        const int ant = k < NMAX ? k : k - NMAX;
        xv[ t ] = ( ant < N && vi ) ? xi[ k < NMAX ? ant : N + ant ] : 0.f;
        zr[ t ] = ( ant < N && vj ) ? zj[ ant ] : 0.f;
        zim[ t ] = ( ant < N && vj ) ? zj[ N + ant ] : 0.f;
    }
    mf32x16 accP = {}, accS = {};
#pragma unroll
    for ( int t = 0; t < K / 2; t++ )
    {
        const bool re = 2 * t + h < NMAX;
        const float zp = re ? zr[ t ] : zim[ t ]; // Zp row = [Re z | Im z]
        const float zs = re ? zim[ t ] : -zr[ t ]; // Zs row = [Im z | -Re z]
        accP = __builtin_amdgcn_mfma_f32_32x32x2f32( zp, xv[ t ], accP, 0, 0, 0 );
        accS = __builtin_amdgcn_mfma_f32_32x32x2f32( zs, xv[ t ], accS, 0, 0, 0 );
    }
    if ( !vi )
        return;
    const float g2 = a.g2, k2 = 2.f * g2;
    const float qi = -g2 * Zn[ a.c0 + i ];
    const int64_t rowBase = gramRow( i );
#pragma unroll
    for ( int v = 0; v < 16; v++ )
    {
        const int64_t j = bj * 32 + 8 * ( v / 4 ) + 4 * h + ( v % 4 );
        if ( j > i || j >= a.nA )
            continue;
        const float q = qi - g2 * Zn[ a.c0 + j ];
        const float p = accP[ v ], s = accS[ v ];
        const float E0 = __builtin_amdgcn_exp2f( fminf( fmaf( k2, p, q ), 0.f ) );
        const float E1 = __builtin_amdgcn_exp2f( fminf( fmaf( k2, s, q ), 0.f ) );
        const float E2 = __builtin_amdgcn_exp2f( fminf( fmaf( -k2, s, q ), 0.f ) );
        float* o = Gact + rowBase + 3 * j;
        o[ 0 ] = a.wL * p + a.wG * E0;
        o[ 1 ] = a.wL * s + a.wG * E1;
        o[ 2 ] = -a.wL * s + a.wG * E2;
    }
}

} // namespace apsm
