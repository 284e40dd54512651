// io.cpp -- MatrixMarket reader and the FP64 host residual check.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "host_common.hpp"
#include "read_matrix.h"

using isai::set_error;

extern "C" {

int isai_read_matrix_market(const char *path, int *nrows, int *ncols, int *nnz, int **rp_out, int **ci_out,
                            double **v_out) {
    ISAI_HOST_REQUIRE(path && nrows && ncols && nnz && rp_out && ci_out && v_out,
                      "isai_read_matrix_market: null argument");
    *rp_out = nullptr;
    *ci_out = nullptr;
    *v_out = nullptr;
    std::ifstream fin(path);
    if (!fin) {
        set_error("Cannot open %s", path);
        return ISAI_ERR_IO;
    }
    std::string line;
    if (!std::getline(fin, line)) {
        set_error("Empty file");
        return ISAI_ERR_IO;
    }
    if (line.find("MatrixMarket") == std::string::npos) {
        set_error("Not a MatrixMarket file");
        return ISAI_ERR_IO;
    }
    do {
        if (!std::getline(fin, line)) {
            set_error("Missing size line");
            return ISAI_ERR_IO;
        }
    } while (!line.empty() && line[0] == '%');
    long long m = 0, n = 0, k = 0;
    {
        std::istringstream iss(line);
        iss >> m >> n >> k;
        if (!iss || m < 0 || n < 0 || k < 0 || m > INT32_MAX || n > INT32_MAX || k > INT32_MAX) {
            set_error("Invalid size line in mtx");
            return ISAI_ERR_IO;
        }
    }
    struct Trip {
        int r, c;
        double v;
    };
    std::vector<Trip> trips;
    trips.reserve((size_t)k);
    for (long long e = 0; e < k; ++e) {
        long long r, c;
        double v;
        fin >> r >> c >> v;
        if (!fin) {
            set_error("Error reading entry %lld", e);
            return ISAI_ERR_IO;
        }
        if (r < 1 || r > m || c < 1 || c > n) {
            set_error("Entry %lld out of range (%lld,%lld)", e, r, c);
            return ISAI_ERR_IO;
        }
        trips.push_back({(int)(r - 1), (int)(c - 1), v});
    }
    std::stable_sort(trips.begin(), trips.end(), [](const Trip &a, const Trip &b) {
        return a.r != b.r ? a.r < b.r : a.c < b.c;
    });
    int *rp = (int *)malloc(sizeof(int) * (size_t)(m + 1));
    int *ci = (int *)malloc(sizeof(int) * (size_t)(k ? k : 1));
    double *v = (double *)malloc(sizeof(double) * (size_t)(k ? k : 1));
    if (!rp || !ci || !v) {
        free(rp);
        free(ci);
        free(v);
        set_error("isai_read_matrix_market: out of host memory");
        return ISAI_ERR_ALLOC;
    }
    std::fill(rp, rp + m + 1, 0);
    for (size_t e = 0; e < trips.size(); ++e) {
        rp[trips[e].r + 1]++;
        ci[e] = trips[e].c;
        v[e] = trips[e].v;
    }
    for (long long r = 0; r < m; ++r) rp[r + 1] += rp[r];
    *nrows = (int)m;
    *ncols = (int)n;
    *nnz = (int)k;
    *rp_out = rp;
    *ci_out = ci;
    *v_out = v;
    return ISAI_OK;
}

int isai_rel_errors_csr(int n, const int *rp, const int *ci, const double *val, const double *b, const double *x_true,
                        const double *x, double *rel_error_x, double *rel_residual) {
    ISAI_HOST_REQUIRE(n >= 0 && rp && ci && val && b && x && rel_residual, "isai_rel_errors_csr: null argument");
    if (rel_error_x) {
        double nd = 0.0, nt = 0.0;
        if (x_true) {
            for (int i = 0; i < n; ++i) {
                const double d = x[i] - x_true[i];
                nd += d * d;
                nt += x_true[i] * x_true[i];
            }
        }
        *rel_error_x = nt > 0.0 ? std::sqrt(nd) / std::sqrt(nt) : 0.0;
    }
    double nr = 0.0, nbv = 0.0;
    for (int i = 0; i < n; ++i) {
        double ax = 0.0;
        for (int p = rp[i]; p < rp[i + 1]; ++p) ax += val[p] * x[ci[p]];
        const double r = b[i] - ax;
        nr += r * r;
        nbv += b[i] * b[i];
    }
    *rel_residual = nbv > 0.0 ? std::sqrt(nr) / std::sqrt(nbv) : INFINITY;
    return ISAI_OK;
}

}  // extern "C"
