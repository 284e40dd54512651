// host_formats.cpp -- host-side format work: CSR->BSR, triangular / BCSC patterns,
// the BCSC->BCSR map, and the synthetic stencil generators.
//
// The reference does all of this on the host with std::vector / std::map
// (csr2bsr.h:14-76, isai_bsr.h:246-284, bcsc2bcsr.h:11-60); here every pass is a
// two-phase count / fill over caller-owned arrays, parallel over block rows.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "bcsc2bcsr.h"
#include "csr2bsr.h"
#include "host_common.hpp"
#include "isai_bsr.h"
#include "isai_gen.h"

using isai::set_error;

namespace {

// Sorted distinct block columns of block row br of a CSR matrix.
void block_cols_of_row(int br, int bs, const int *rp, const int *ci, std::vector<int> &out) {
    out.clear();
    for (int p = rp[br * bs]; p < rp[(br + 1) * bs]; ++p) out.push_back(ci[p] / bs);
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- csr2bsr.h
int isai_csr_to_bsr_count(int n, int bs, const int *rp, const int *ci, int *nb, int *nnzb) {
    ISAI_HOST_REQUIRE(rp && nb && nnzb, "isai_csr_to_bsr_count: null argument");
    ISAI_HOST_REQUIRE(bs >= 1 && bs <= 8, "isai_csr_to_bsr_count: bs=%d outside 1..8", bs);
    ISAI_HOST_REQUIRE(n >= 0 && n % bs == 0,
                      "csr_to_bsr: n = %d is not a multiple of block size %d", n, bs);
    const int nbl = n / bs;
    long long total = 0;
#pragma omp parallel reduction(+ : total)
    {
        std::vector<int> cols;
#pragma omp for schedule(static)
        for (int br = 0; br < nbl; ++br) {
            block_cols_of_row(br, bs, rp, ci, cols);
            total += (long long)cols.size();
        }
    }
    if (total > 0x7fffffffLL) {
        set_error("csr_to_bsr: nnzb %lld exceeds int32", total);
        return ISAI_ERR_CAPACITY;
    }
    *nb = nbl;
    *nnzb = (int)total;
    return ISAI_OK;
}

int isai_csr_to_bsr_fill(int n, int bs, const int *rp, const int *ci, const double *v,
                         int *brp, int *bci, double *bv) {
    ISAI_HOST_REQUIRE(rp && brp && bci && bv, "isai_csr_to_bsr_fill: null argument");
    ISAI_HOST_REQUIRE(bs >= 1 && bs <= 8, "isai_csr_to_bsr_fill: bs=%d outside 1..8", bs);
    ISAI_HOST_REQUIRE(n >= 0 && n % bs == 0,
                      "csr_to_bsr: n = %d is not a multiple of block size %d", n, bs);
    const int nbl = n / bs, bs2 = bs * bs;
    // Pass 1: block counts per row, then the row pointer.
    std::vector<int> cnt(nbl + 1, 0);
#pragma omp parallel
    {
        std::vector<int> cols;
#pragma omp for schedule(static)
        for (int br = 0; br < nbl; ++br) {
            block_cols_of_row(br, bs, rp, ci, cols);
            cnt[br] = (int)cols.size();
        }
    }
    brp[0] = 0;
    for (int br = 0; br < nbl; ++br) brp[br + 1] = brp[br] + cnt[br];
    // Pass 2: columns, then values in CSR order; the first entry of a block is
    // assigned into the zero-filled block and later ones are added, which keeps
    // a lone -0.0 entry as -0.0 exactly as the reference's std::map path does.
#pragma omp parallel
    {
        std::vector<int> cols;
        std::vector<char> seen;
#pragma omp for schedule(static)
        for (int br = 0; br < nbl; ++br) {
            block_cols_of_row(br, bs, rp, ci, cols);
            const int base = brp[br];
            const int u = (int)cols.size();
            seen.assign(u, 0);
            for (int q = 0; q < u; ++q) {
                bci[base + q] = cols[q];
                std::memset(bv + (size_t)(base + q) * bs2, 0, sizeof(double) * bs2);
            }
            for (int lr = 0; lr < bs; ++lr) {
                const int row = br * bs + lr;
                for (int p = rp[row]; p < rp[row + 1]; ++p) {
                    const int bc = ci[p] / bs, lc = ci[p] % bs;
                    const int q = (int)(std::lower_bound(cols.begin(), cols.end(), bc) - cols.begin());
                    double *blk = bv + (size_t)(base + q) * bs2;
                    if (!seen[q]) {
                        blk[lr * bs + lc] = v[p];
                        seen[q] = 1;
                    } else {
                        blk[lr * bs + lc] += v[p];
                    }
                }
            }
        }
    }
    return ISAI_OK;
}

// -------------------------------------------------------------- bcsc2bcsr.h
int isai_tri_bcsr_pattern_count(int nb, const int *Arow, const int *Acol, int lower, int *nnz) {
    ISAI_HOST_REQUIRE(Arow && Acol && nnz && nb >= 0, "isai_tri_bcsr_pattern_count: bad argument");
    long long k = 0;
    for (int r = 0; r < nb; ++r)
        for (int p = Arow[r]; p < Arow[r + 1]; ++p) {
            const int c = Acol[p];
            if (lower ? (c <= r) : (c >= r)) ++k;
        }
    *nnz = (int)k;
    return ISAI_OK;
}

int isai_tri_bcsr_pattern(int nb, const int *Arow, const int *Acol, int lower, int *Trow, int *Tcol) {
    ISAI_HOST_REQUIRE(Arow && Acol && Trow && Tcol && nb >= 0, "isai_tri_bcsr_pattern: bad argument");
    int k = 0;
    Trow[0] = 0;
    for (int r = 0; r < nb; ++r) {
        for (int p = Arow[r]; p < Arow[r + 1]; ++p) {
            const int c = Acol[p];
            if (lower ? (c <= r) : (c >= r)) Tcol[k++] = c;
        }
        Trow[r + 1] = k;
    }
    return ISAI_OK;
}

int isai_map_bcsc2bcsr(int nb, const int *colptr, const int *rowind, const int *rowptr,
                       const int *colind, int *map) {
    ISAI_HOST_REQUIRE(colptr && rowind && rowptr && colind && map && nb >= 0,
                      "isai_map_bcsc2bcsr: bad argument");
    int bad_r = -1, bad_c = -1;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nb; ++c) {
        for (int k = colptr[c]; k < colptr[c + 1]; ++k) {
            const int r = rowind[k];
            int found = -1;
            for (int p = rowptr[r]; p < rowptr[r + 1]; ++p)
                if (colind[p] == c) { found = p; break; }
            map[k] = found;
            if (found < 0) {
#pragma omp critical
                { if (bad_r < 0) { bad_r = r; bad_c = c; } }
            }
        }
    }
    if (bad_r >= 0) {
        set_error("Pattern mismatch at (%d,%d)", bad_r, bad_c);
        return ISAI_ERR_PATTERN;
    }
    return ISAI_OK;
}

// ---------------------------------------------------------------- isai_bsr.h
int isai_bcsc_pattern_count(int nb, const int *Arow, const int *Acol, int lower, int *nnz) {
    ISAI_HOST_REQUIRE(Arow && Acol && nnz && nb >= 0, "isai_bcsc_pattern_count: bad argument");
    long long k = 0;
    for (int br = 0; br < nb; ++br) {
        bool diag = false;
        for (int p = Arow[br]; p < Arow[br + 1]; ++p) {
            const int bc = Acol[p];
            if (lower ? (bc > br) : (bc < br)) continue;
            if (bc == br) diag = true;
            ++k;
        }
        if (!diag) ++k;
    }
    *nnz = (int)k;
    return ISAI_OK;
}

int isai_bcsc_pattern(int nb, const int *Arow, const int *Acol, int lower, int *Mbrow, int *Mcol) {
    ISAI_HOST_REQUIRE(Arow && Acol && Mbrow && Mcol && nb >= 0, "isai_bcsc_pattern: bad argument");
    // Transpose of the kept triangle with the diagonal forced: column j lists the
    // rows br with block (br, j) kept, ascending and without duplicates
    // (isai_bsr.h:257-283).  Rows are visited in ascending order, so each column
    // fills in ascending order; a final per-column sort covers unsorted inputs.
    std::vector<int> cnt(nb + 1, 0);
    for (int br = 0; br < nb; ++br) {
        bool diag = false;
        for (int p = Arow[br]; p < Arow[br + 1]; ++p) {
            const int bc = Acol[p];
            if (lower ? (bc > br) : (bc < br)) continue;
            if (bc == br) diag = true;
            ++cnt[bc];
        }
        if (!diag) ++cnt[br];
    }
    Mbrow[0] = 0;
    for (int j = 0; j < nb; ++j) Mbrow[j + 1] = Mbrow[j] + cnt[j];
    std::vector<int> fill(Mbrow, Mbrow + nb);
    for (int br = 0; br < nb; ++br) {
        bool diag = false;
        for (int p = Arow[br]; p < Arow[br + 1]; ++p) {
            const int bc = Acol[p];
            if (lower ? (bc > br) : (bc < br)) continue;
            if (bc == br) diag = true;
            Mcol[fill[bc]++] = br;
        }
        if (!diag) Mcol[fill[br]++] = br;
    }
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nb; ++j) {
        int *b = Mcol + Mbrow[j], *e = Mcol + Mbrow[j + 1];
        if (!std::is_sorted(b, e)) std::sort(b, e);
    }
    return ISAI_OK;
}

// ------------------------------------------------------------------ isai_gen.h
namespace {

struct Rng {
    std::mt19937_64 g;
    explicit Rng(uint64_t seed) : g(seed) {}
    double u01() { return (double)(g() >> 11) * 0x1.0p-53; }
    double uni(double a, double b) { return a + (b - a) * u01(); }
};

struct Offset { int dx, dy, dz; };

std::vector<Offset> stencil(int problem) {
    std::vector<Offset> o;
    if (problem == 0) {  // 2-D 5-point, ascending linear offset
        o = {{0, -1, 0}, {-1, 0, 0}, {0, 0, 0}, {1, 0, 0}, {0, 1, 0}};
    } else if (problem == 1) {  // 3-D 7-point
        o = {{0, 0, -1}, {0, -1, 0}, {-1, 0, 0}, {0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    } else if (problem == 2) {  // 3-D 27-point
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) o.push_back({dx, dy, dz});
    } else {  // 2-D 9-point (problems 3, 4)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) o.push_back({dx, dy, 0});
    }
    return o;
}

bool is3d(int problem) { return problem == 1 || problem == 2; }

int default_L(int problem) {
    switch (problem) {
        case 0: return 64;
        case 1: return 96;
        case 2: return 64;
        case 4: return 256;
        default: return 0;
    }
}

}  // namespace

int isai_gen_problem_count(int problem, int L, int *nb, int *nnzb) {
    ISAI_HOST_REQUIRE(nb && nnzb, "isai_gen_problem_count: null output");
    ISAI_HOST_REQUIRE(problem >= 0 && problem <= 4, "isai_gen: unknown problem %d", problem);
    if (L <= 0) L = default_L(problem);
    ISAI_HOST_REQUIRE(L >= 1, "isai_gen: problem %d needs an explicit L", problem);
    const int Lz = is3d(problem) ? L : 1;
    const long long nodes = (long long)L * L * Lz;
    ISAI_HOST_REQUIRE(nodes < (1LL << 28), "isai_gen: grid too large");
    long long k = 0;
    const auto off = stencil(problem);
    for (int z = 0; z < Lz; ++z)
        for (int y = 0; y < L; ++y)
            for (int x = 0; x < L; ++x)
                for (const auto &o : off) {
                    const int xx = x + o.dx, yy = y + o.dy, zz = z + o.dz;
                    if (xx >= 0 && xx < L && yy >= 0 && yy < L && zz >= 0 && zz < Lz) ++k;
                }
    ISAI_HOST_REQUIRE(k < (1LL << 31), "isai_gen: nnzb overflows int32");
    *nb = (int)nodes;
    *nnzb = (int)k;
    return ISAI_OK;
}

int isai_gen_problem_fill(int problem, int L, uint64_t seed, int *rp, int *ci, double *val) {
    ISAI_HOST_REQUIRE(rp && ci && val, "isai_gen_problem_fill: null output");
    ISAI_HOST_REQUIRE(problem >= 0 && problem <= 4, "isai_gen: unknown problem %d", problem);
    if (L <= 0) L = default_L(problem);
    ISAI_HOST_REQUIRE(L >= 1, "isai_gen: problem %d needs an explicit L", problem);
    const int Lz = is3d(problem) ? L : 1;
    const auto off = stencil(problem);
    Rng rng(seed);
    // Upwinded convection for problem 2: velocity (1, 0.5, 0.25) scaled to cell
    // Peclet 0.5 against unit diffusion.
    const double vx = 0.5 * 1.0, vy = 0.5 * 0.5, vz = 0.5 * 0.25;
    int k = 0;
    rp[0] = 0;
    int row = 0;
    for (int z = 0; z < Lz; ++z)
        for (int y = 0; y < L; ++y)
            for (int x = 0; x < L; ++x, ++row) {
                for (const auto &o : off) {
                    const int xx = x + o.dx, yy = y + o.dy, zz = z + o.dz;
                    if (xx < 0 || xx >= L || yy < 0 || yy >= L || zz < 0 || zz >= Lz) continue;
                    const bool diag = (o.dx == 0 && o.dy == 0 && o.dz == 0);
                    ci[k] = xx + L * (yy + L * zz);
                    double *b = val + (size_t)k * 16;
                    if (problem == 0 || problem == 4) {
                        if (diag) {
                            double R[16];
                            for (int e = 0; e < 16; ++e) R[e] = rng.uni(-0.1, 0.1);
                            for (int r = 0; r < 4; ++r)
                                for (int c = 0; c < 4; ++c)
                                    b[r * 4 + c] = (r == c ? 4.5 : 0.0) + 0.5 * (R[r * 4 + c] + R[c * 4 + r]);
                        } else {
                            const double u = rng.uni(0.9, 1.1);
                            for (int e = 0; e < 16; ++e) b[e] = (e % 5 == 0) ? -u : 0.0;
                        }
                    } else if (problem == 1) {
                        for (int e = 0; e < 16; ++e) {
                            const double d = (e % 5 == 0) ? 1.0 : 0.0;
                            b[e] = diag ? 6.5 * d + rng.uni(-0.1, 0.1) : -d + rng.uni(-0.05, 0.05);
                        }
                    } else if (problem == 2) {
                        double a = diag ? 26.0 + vx + vy + vz : -1.0;
                        if (o.dx == -1 && o.dy == 0 && o.dz == 0) a -= vx;
                        if (o.dx == 0 && o.dy == -1 && o.dz == 0) a -= vy;
                        if (o.dx == 0 && o.dy == 0 && o.dz == -1) a -= vz;
                        double S[16];
                        for (int r = 0; r < 4; ++r)
                            for (int c = r; c < 4; ++c) {
                                S[r * 4 + c] = rng.uni(-0.05, 0.05);
                                S[c * 4 + r] = S[r * 4 + c];
                            }
                        for (int e = 0; e < 16; ++e) b[e] = a * (((e % 5) == 0 ? 1.0 : 0.0) + S[e]);
                    } else {  // problem 3
                        if (diag) {
                            for (int e = 0; e < 16; ++e) b[e] = ((e % 5) == 0 ? 8.5 : 0.0) + rng.uni(-0.2, 0.2);
                        } else {
                            const double u = rng.uni(0.5, 1.5);
                            for (int e = 0; e < 16; ++e) b[e] = (e % 5 == 0) ? -u : 0.0;
                        }
                    }
                    ++k;
                }
                rp[row + 1] = k;
            }
    return ISAI_OK;
}

int isai_gen_normal(uint64_t seed, size_t n, double *out) {
    ISAI_HOST_REQUIRE(out || n == 0, "isai_gen_normal: null output");
    Rng rng(seed);
    const double two_pi = 6.283185307179586476925286766559;
    for (size_t i = 0; i < n; i += 2) {
        const double u1 = rng.u01(), u2 = rng.u01();
        const double r = std::sqrt(-2.0 * std::log(1.0 - u1));
        out[i] = r * std::cos(two_pi * u2);
        if (i + 1 < n) out[i + 1] = r * std::sin(two_pi * u2);
    }
    return ISAI_OK;
}

}  // extern "C"
