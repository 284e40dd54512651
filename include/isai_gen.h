/*
 * isai_gen.h -- synthetic block-stencil problems (bs = 4) in BSR form.
 *
 * No reference counterpart: the reference reads MatrixMarket files
 * (read_matrix.h:12-99, main.cu:478-488) and ships none.  These generators make
 * the benchmark workloads reproducible from (problem, L, seed).
 *
 * Node order is lexicographic with x fastest (id = x + L*y + L*L*z); one 4x4 block
 * per (node, neighbour); block columns ascending.  Randomness: std::mt19937_64(seed),
 * u01 = (next() >> 11) * 2^-53, drawn block row by block row, blocks in ascending
 * column order, entries row-major within a block.
 *
 *  problem 0  2-D 5-point,  D = 4.5 I + (R + R^T)/2, R ~ U(-0.1,0.1) (16 draws);
 *             off-diagonal -u I, u ~ U(0.9,1.1) (1 draw).            default L = 64
 *  problem 1  3-D 7-point,  D = 6.5 I + U(-0.1,0.1) (16 draws);
 *             off-diagonal -I + U(-0.05,0.05) (16 draws).            default L = 96
 *  problem 2  3-D 27-point upwinded convection-diffusion: scalar stencil 26 / -1
 *             (diffusion 1), velocity (1,0.5,0.25) at cell Peclet 0.5 added
 *             upwind (+v to the centre, -v to the x-1/y-1/z-1 neighbour); each
 *             block = a_IJ (I + S), S symmetric, upper triangle U(-0.05,0.05)
 *             (10 draws, row-major over c >= r).                      default L = 64
 *  problem 3  2-D 9-point,  D = 8.5 I + U(-0.2,0.2) (16 draws);
 *             off-diagonal -u I, u ~ U(0.5,1.5) (1 draw).             L given per system
 *  problem 4  2-D 9-point with problem 0's coefficients.            default L = 256
 */
#ifndef ISAI_GEN_H
#define ISAI_GEN_H

#include "isai_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* L <= 0 selects the problem's default grid size. */
int isai_gen_problem_count(int problem, int L, int *nb, int *nnzb);
int isai_gen_problem_fill(int problem, int L, uint64_t seed,
                          int *bsrRowPtr, int *bsrColInd, double *bsrVal);
/* n doubles, N(0,1) by Box-Muller over the same u01 stream (seeded separately). */
int isai_gen_normal(uint64_t seed, size_t n, double *out);

#ifdef __cplusplus
}
#endif
#endif
