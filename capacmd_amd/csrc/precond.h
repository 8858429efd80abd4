// Shared pieces of the block-Jacobi preconditioner (precond.hip) and the
// GMRES kernels that apply it (solver.hip).
//
// Block b of the preconditioner is metal atoms 64 b .. 64 b + 63 (the relay's
// tile blocks, relay_tile.h); its 256 unknowns rho are
//     rho < 192: dipole component rho % 3 of atom 64 b + rho / 3
//     rho >= 192: charge of atom 64 b + rho - 192
// Vector launches aligned to the blocks take four workgroups per block, one
// per quarter of its unknowns (quarter 3: the charges), numbered so that the
// four share an XCD (workgroups are dealt to the 8 XCDs round robin): g =
// 32 (b / 8) + 8 s + b % 8 for quarter s of block b.
#pragma once

#include "capac_internal.h"
#include "relay_tile.h"

constexpr int kBjUnk = 4 * kTileEdge;          // unknowns per block
constexpr int kCgMax = kBjUnk;                  // coarse unknowns at most (4 per block: 4096 metal atoms)

// vector row of block-local unknown rho of block b, or -1 past Nm
__device__ __forceinline__ int bj_row(int b, int rho, int Nm) {
    const int atom = b * kTileEdge + (rho < 3 * kTileEdge ? rho / 3 : rho - 3 * kTileEdge);
    if (atom >= Nm) return -1;
    return rho < 3 * kTileEdge ? 3 * kTileEdge * b + rho : 3 * Nm + atom;
}

// workgroups of a block-aligned launch over Nb blocks (some past Nb idle)
__host__ __device__ constexpr int bj_groups(int Nb) { return 32 * ((Nb + 7) / 8); }

// block and quarter of workgroup g
__device__ __forceinline__ void bj_group(int g, int& b, int& s) {
    const int r = g >> 3;
    s = r & 3;
    b = (r >> 2) * 8 + (g & 7);
}
