// sptrsv.hip -- sync-free triangular solves with the packed ILU(0) factor on gfx950
// (replaces spTrSolveL / spTrSolveU / csr_cape_SpTRSV, reference cape_sptrsv.h:8-179).
//
// The reference gives every scalar row one thread that spins on a volatile bool
// flag of each dependency and relies on independent thread scheduling inside a
// warp (cape_sptrsv.h:36-51).  A wave64 has no independent thread scheduling, so
// a lane spinning on a value another lane of its own wave must produce would
// never finish.  Here, in two schedules of the BSR form:
//   * chunk schedule: one wave owns a CHUNK of consecutive rows (ascending for L,
//     descending for U) taken from an atomic ticket, and walks them in order:
//     dependencies inside the chunk are read from LDS, never waited on;
//   * row schedule: one quad of lanes per block row, and the quads of a wave
//     advance independently of each other; rows in level order, or in unit order
//     (units of 64 rows grown so that most dependencies stay inside, exchanged
//     through LDS, with one wave per workgroup importing the rest);
//   * dependencies on other chunks (or rows) are waited on with per-row ready flags that
//     live in the solution vector itself: it is pre-filled with a NaN sentinel
//     and every value is published with an agent-scope atomic store, so a reader
//     polling with agent-scope atomic loads sees either the sentinel or the final
//     value -- no fence, no separate flag array, no host synchronisation and no
//     grid-wide barrier;
//   * tickets are handed out in row order, so a wave only ever waits on chunks
//     already held by running waves (forward progress under any dispatch order);
//   * every wait is bounded; an exhausted wait sets an error word instead of
//     hanging, which the host reads with isai_sptrsv_status().
// The last wave of a solve resets the ticket counters, so consecutive solves
// (or graph replays) need no memset.
#include <algorithm>
#include <atomic>
#include <climits>
#include <thread>
#include <vector>

#include "cape_sptrsv.h"
#include "common.hpp"
#include "internal.hpp"
#include "isai_bsr.h"

namespace isai {

constexpr unsigned long long kSentinel = 0x7FF4DEADBEEF0001ull;  // signalling-NaN payload
constexpr int kMaxChunk = 256;
constexpr int kCsrMaxChunk = 1024;

struct Counters {
    unsigned int ticket_l, done_l, ticket_u, done_u;
    int err;
    int spin_limit;
    int pad[2];
};

__device__ __forceinline__ double poll(const double *p, int limit, bool &timed_out) {
    unsigned long long *q = (unsigned long long *)(p);
    unsigned long long v = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int spins = 0;
    while (v == kSentinel) {
        if (++spins > limit) {
            timed_out = true;
            break;
        }
        __builtin_amdgcn_s_sleep(1);
        v = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return __longlong_as_double((long long)v);
}

// Value of lane K of the quad (lanes 0..3), in every lane of the quad: a DPP
// quad_perm [K,K,K,K] move of both halves, no LDS round trip.
template <int K>
__device__ __forceinline__ double quad_bcast(double v) {
    constexpr int ctrl = K | (K << 2) | (K << 4) | (K << 6);
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(u & 0xffffffffll), ctrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), ctrl, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ void publish(double *p, double v) {
    __hip_atomic_store((unsigned long long *)p, (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// -------------------------------------------------------------- BSR, bs = 4
// Rows are handled in "processing positions" q: q = i for L (ascending), q = nb-1-i
// for U (descending), so both triangles share one kernel.  A chunk is a run of
// positions [qlo, qhi) whose boundaries follow the dependency chains.  The factor
// values used here are pre-scaled: S_ik = inv(D_i) T_ik for every off-diagonal
// block (T = L or U, D_i its diagonal part), so row i of the solve is
//     x_i = inv(D_i) rhs_i - sum_k S_ik x_k .
// A dependency k of row i is "local" when it lies in the same chunk (its x is in
// LDS) and "remote" otherwise (an item polled from the solution vector).
//
// One wave per chunk:
//   prologue  stage zb = inv(D_i) rhs_i, the local blocks S_ik and the per-row
//             item tables of the chunk in LDS;
//   loop      (a) remote phase: each of the 64 lanes takes one not-yet-done remote
//                 item from a window starting at the chain's current row, reads
//                 its x_k with ONE agent-scope atomic poll (no spinning) and, if
//                 published, stores S_ik x_k in the item's LDS slot;
//             (b) chain phase: four lanes (one per row of the 4x4 block) advance
//                 the chunk's rows in order while all remote items of the next row
//                 are in, touching only LDS, and publish each x_i.
// Every step either makes progress or counts toward the stall limit, so the
// kernel always terminates; remote items only ever point at chunks with lower
// tickets, which are already running.
constexpr int kPipeRows = 128;   // rows per chunk
constexpr int kPipeRemote = 512; // remote items per chunk
constexpr int kPipeLocal = 128;  // local items per chunk

static size_t pipe_lds_bytes(int rows, int rem, int loc) {
    // at least one remote item slot, so that the chain's clamped reads stay inside
    rem = std::max(rem, 1);
    return sizeof(double) * ((size_t)rows * 8 + (size_t)rem * 4 + (size_t)loc * 16) +
           sizeof(int) * ((size_t)rem * 3 + 2 * (size_t)loc + 2 * ((size_t)rows + 1));
}

template <bool UPPER>
__global__ __launch_bounds__(64) void bsr4_sptrsv_pipe_kernel(
    int nb, const int4 *__restrict__ chunks, const int *__restrict__ ci, const double *__restrict__ S,
    const double *__restrict__ dinv, const int *__restrict__ rem_ptr, const int *__restrict__ rem_items,
    const int *__restrict__ loc_ptr, const int *__restrict__ loc_items, const double *rhs, double *out,
    double *sentinel_fill, Counters *cnt, const int *__restrict__ skip, int max_rows, int max_rem, int max_loc,
    int poll_min) {
    if (skip && *skip) return;
    // LDS carved to the largest chunk of this triangle
    extern __shared__ double dyn[];
    max_rem = max(max_rem, 1);
    double *zb = dyn;                         // max_rows * 4
    double *xs = zb + max_rows * 4;           // max_rows * 4
    double *rem = xs + max_rows * 4;          // max_rem * 4
    double *locv = rem + max_rem * 4;         // max_loc * 16
    int *done = (int *)(locv + max_loc * 16); // max_rem
    int *ik = done + max_rem;                 // max_rem: column position of the item's block
    int *ip = ik + max_rem;                   // max_rem: block position of the item
    int *lk = ip + max_rem;                   // max_loc
    int *rrp = lk + max_loc;                  // max_rows + 1
    int *lrp = rrp + max_rows + 1;            // max_rows + 1
    int *lpos = lrp + max_rows + 1;           // max_loc: block position of the local item
    __shared__ int s_c, s_chain;
    const int lane = threadIdx.x;
    unsigned int *ticket = UPPER ? &cnt->ticket_u : &cnt->ticket_l;
    unsigned int *donec = UPPER ? &cnt->done_u : &cnt->done_l;
    const int limit = cnt->spin_limit;
    if (lane == 0) s_c = (int)atomicAdd(ticket, 1u);
    __syncthreads();
    // chunk c: [2c] = (qlo, qhi, first remote item, first local item), [2c+1] = (end remote, end local, level, 0)
    const int4 ch = chunks[2 * s_c];
    const int4 ce = chunks[2 * s_c + 1];
    const int qlo = ch.x, qhi = ch.y, r0 = ch.z, l0 = ch.w;
    const int nrows = qhi - qlo, nrem = ce.x - r0, nloc = ce.y - l0;
    auto row_of = [&](int q) { return UPPER ? nb - 1 - q : q; };

    // ---- prologue.  The loops are unrolled so that several iterations' loads
    // are in flight together: a chunk holds its residency slot while it stages,
    // and one dependent load chain per iteration took a quarter of its life.
    const int nz = nrows * 4;
#pragma unroll 4
    for (int t = lane; t < nz; t += 64) {
        const int rr = t >> 2, m = t & 3;
        const int i = row_of(qlo + rr);
        const double *D = dinv + (size_t)i * 16 + m * 4;
        const double *b = rhs + (size_t)i * 4;
        zb[t] = D[0] * b[0] + D[1] * b[1] + D[2] * b[2] + D[3] * b[3];
    }
#pragma unroll 2
    for (int t = lane; t <= nrows; t += 64) {
        rrp[t] = rem_ptr[qlo + t] - r0;
        lrp[t] = loc_ptr[qlo + t] - l0;
    }
#pragma unroll 2
    for (int t = lane; t < nloc; t += 64) {
        const int p = loc_items[l0 + t];
        const int k = ci[p];
        lpos[t] = p;
        lk[t] = (UPPER ? nb - 1 - k : k) - qlo;
    }
#pragma unroll 4
    for (int t = lane; t < nrem; t += 64) {
        const int p = rem_items[r0 + t];
        ip[t] = p;
        ik[t] = ci[p];
        done[t] = 0;
    }
    __syncthreads();
    // the local blocks: positions from LDS, all loads of the loop in flight
#pragma unroll 4
    for (int t = lane; t < nloc * 16; t += 64) locv[t] = S[(size_t)lpos[t >> 4] * 16 + (t & 15)];
    __syncthreads();

    int chain = 0, stall = 0;
    bool timed_out = false;
    // The S block of the item a lane polls is fetched once, when the item enters
    // the lane's window (usually several rounds before its x is published), and
    // kept in registers: a published x then costs one poll, not a poll followed
    // by a dependent load of the block.
    int held = -1;
    v2d b0 = {0.0, 0.0}, b1 = b0, b2 = b0, b3 = b0, b4 = b0, b5 = b0, b6 = b0, b7 = b0;
    // chain lanes: x of the row solved last (row xp_row of the chunk), all four components
    int xp_row = -1;
    double xp0 = 0.0, xp1 = 0.0, xp2 = 0.0, xp3 = 0.0;
    // Polls are write-through loads served beyond L2, and every resident chunk
    // polls at once, so the window adapts: a round polls about twice as many items
    // as the previous round found published (at least kPollMin), instead of always
    // 64 items up to 32 rows ahead of what the producers have reached.
    int window = 64;
    while (chain < nrows) {
        // ---- (a) remote phase: one poll per pending item of the window
        int got = 0;
        const int w = rrp[chain] + lane;
        if (lane < window && w < nrem && !done[w]) {
            const int k = ik[w];
            unsigned long long *xq = (unsigned long long *)(out + (size_t)k * 4);
            const unsigned long long u0 = __hip_atomic_load(xq + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned long long u1 = __hip_atomic_load(xq + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned long long u2 = __hip_atomic_load(xq + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned long long u3 = __hip_atomic_load(xq + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (w != held) {
                const v2d *B = (const v2d *)(S + (size_t)ip[w] * 16);
                b0 = B[0], b1 = B[1], b2 = B[2], b3 = B[3], b4 = B[4], b5 = B[5], b6 = B[6], b7 = B[7];
                held = w;
            }
            if (u0 != kSentinel && u1 != kSentinel && u2 != kSentinel && u3 != kSentinel) {
                const double x0 = __longlong_as_double((long long)u0), x1 = __longlong_as_double((long long)u1);
                const double x2 = __longlong_as_double((long long)u2), x3 = __longlong_as_double((long long)u3);
                rem[w * 4 + 0] = b0.x * x0 + b0.y * x1 + b1.x * x2 + b1.y * x3;
                rem[w * 4 + 1] = b2.x * x0 + b2.y * x1 + b3.x * x2 + b3.y * x3;
                rem[w * 4 + 2] = b4.x * x0 + b4.y * x1 + b5.x * x2 + b5.y * x3;
                rem[w * 4 + 3] = b6.x * x0 + b6.y * x1 + b7.x * x2 + b7.y * x3;
                done[w] = 1;
                got = 1;
            }
        }
        __syncthreads();
        // ---- (b) chain phase: lanes 0..3 advance while the next row is complete
        // Rows with at most 2 remote and 1 local item (stencil rows) take a fast
        // path: all of the row's LDS reads are issued before any is used (indices
        // clamped into the arrays, results selected afterwards), and the local item
        // of the row solved just before reads its x from registers (quad broadcast)
        // instead of LDS.  The arithmetic and its order are those of the loop path.
        if (lane < 4) {
            int adv = 0;
            while (chain < nrows && adv < 64) {
                const int ra = rrp[chain], rb = rrp[chain + 1];
                const int la = lrp[chain], lb = lrp[chain + 1];
                double z;
                if (rb - ra <= 2 && lb - la <= 1) {
                    const int r1 = min(ra + 1, nrem - 1), r0c = min(ra, nrem - 1);
                    const int l0c = min(la, max(nloc - 1, 0));
                    const int d0 = done[r0c], d1 = done[r1];
                    const double v0 = rem[r0c * 4 + lane], v1 = rem[r1 * 4 + lane];
                    const double zc = zb[chain * 4 + lane];
                    const int kk = lk[l0c];
                    const double s0 = locv[l0c * 16 + lane * 4 + 0], s1 = locv[l0c * 16 + lane * 4 + 1];
                    const double s2 = locv[l0c * 16 + lane * 4 + 2], s3 = locv[l0c * 16 + lane * 4 + 3];
                    const bool ready = (ra >= rb || d0) && (ra + 1 >= rb || d1);
                    if (!ready) break;
                    z = zc;
                    if (ra < rb) z -= v0;
                    if (ra + 1 < rb) z -= v1;
                    if (la < lb) {
                        double x0, x1, x2, x3;
                        if (kk == xp_row) {
                            x0 = xp0, x1 = xp1, x2 = xp2, x3 = xp3;
                        } else {
                            const double *xk = xs + kk * 4;
                            x0 = xk[0], x1 = xk[1], x2 = xk[2], x3 = xk[3];
                        }
                        z -= s0 * x0 + s1 * x1 + s2 * x2 + s3 * x3;
                    }
                } else {
                    bool ready = true;
                    for (int t = ra; t < rb; ++t) ready = ready && done[t];
                    if (!ready) break;
                    z = zb[chain * 4 + lane];
                    for (int t = ra; t < rb; ++t) z -= rem[t * 4 + lane];
                    for (int t = la; t < lb; ++t) {
                        const double *Sv = locv + t * 16 + lane * 4;
                        const double *xk = xs + lk[t] * 4;
                        z -= Sv[0] * xk[0] + Sv[1] * xk[1] + Sv[2] * xk[2] + Sv[3] * xk[3];
                    }
                }
                xs[chain * 4 + lane] = z;
                xp0 = quad_bcast<0>(z);
                xp1 = quad_bcast<1>(z);
                xp2 = quad_bcast<2>(z);
                xp3 = quad_bcast<3>(z);
                xp_row = chain;
                const int i = row_of(qlo + chain);
                publish(out + (size_t)i * 4 + lane, z);
                if (sentinel_fill) sentinel_fill[(size_t)i * 4 + lane] = __longlong_as_double((long long)kSentinel);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                ++chain;
                ++adv;
            }
            if (lane == 0) s_chain = chain | (adv > 0 ? 0x40000000 : 0);
        }
        __syncthreads();
        const int sc = s_chain;
        if (poll_min < 64) {
            const int found = __popcll(__ballot(got));
            window = min(64, max(poll_min, 2 * found + poll_min));
        }
        const bool progressed = (sc & 0x40000000) || __any(got);
        chain = sc & 0x3fffffff;
        stall = progressed ? 0 : stall + 1;
        if (stall > limit) {
            // give up on the missing items: their contributions read as NaN
            timed_out = true;
            for (int t = lane; t < nrem; t += 64)
                if (!done[t]) {
                    rem[t * 4 + 0] = rem[t * 4 + 1] = rem[t * 4 + 2] = rem[t * 4 + 3] = __longlong_as_double((long long)kSentinel);
                    done[t] = 1;
                }
            stall = 0;
            __syncthreads();
        }
        __builtin_amdgcn_s_sleep(0);
    }
    if (timed_out && lane == 0) atomicOr(&cnt->err, 1);
    if (lane == 0) {
        if (atomicAdd(donec, 1u) == (unsigned)gridDim.x - 1) {
            atomicExch(ticket, 0u);
            atomicExch(donec, 0u);
        }
    }
}

// ------------------------------------------------ BSR, bs = 4, row schedule
// One quad (four lanes, one per row of the 4x4 block) per block row; the rows are
// taken in the plan's order -- level order (a row's level is one more than the
// deepest level among its dependencies) or unit order -- 64 rows per workgroup
// from an atomic ticket.  Both orders place a row after all of its dependencies,
// so they have lower positions, hence a lower or equal ticket: their workgroups
// are resident.  Within a wave the quads advance
// independently of each other: a quad whose inputs are all published computes
// and publishes its row in the same pass in which the other quads keep polling,
// so rows of a wave may depend on each other.
//
// Per pass a quad holds up to kRowBatch dependencies in registers (each lane
// holds row m of every S block and polls component m of every x_k); with all of
// them published the quad subtracts their products, in the row's column order,
// and moves on to the next batch or publishes.
constexpr int kRowWG = 256;  // threads per workgroup
constexpr int kRowsPerWG = kRowWG / 4;
constexpr int kRowBatch = 8;
// Residency.  Every resident row polls, and the polls of all of them share the
// fabric beyond L2: at four workgroups per CU (the register limit) the 96^3
// 7-point factor's publish-to-observe hop took 2.4 us, at one 2.0 us, and the
// solve ran 16 % faster.  Unused dynamic LDS of 96 KB per workgroup keeps one
// workgroup (64 rows) per CU; its start-up stays off the critical path.
constexpr int kRowLds = 96 * 1024;
// The same kernel in unit order (below) holds rows longer -- a unit spans about
// nine levels -- and needs more rows resident: three workgroups per CU.
constexpr int kUnitRowLds = 48 * 1024;
// The row schedule is chosen by default when the triangles' levels average at
// least this many rows; narrower level structures (long chains) keep the chunk
// schedule, whose chains resolve inside a wave.
constexpr int kRowMinWidth = 16;

// AND of a predicate over the four lanes of each quad (DPP quad_perm swaps).
__device__ __forceinline__ bool quad_all(bool v) {
    int a = v ? 1 : 0;
    a &= __builtin_amdgcn_mov_dpp(a, 0xB1, 0xF, 0xF, false);  // [1,0,3,2]
    a &= __builtin_amdgcn_mov_dpp(a, 0x4E, 0xF, 0xF, false);  // [2,3,0,1]
    return a != 0;
}

// info[q] = (row, first dependency, end of dependencies, 1) for position q of
// the processing order.  cx[p] is the column of dependency block p, or ~s when
// that column is the row at slot s (position % 64) of the same workgroup: such a
// value is exchanged through LDS (a row_x slot, pre-filled with the sentinel)
// instead of the round trip through L2 that an agent-scope store and poll cost.
template <bool UPPER>
__global__ __launch_bounds__(kRowWG) void bsr4_sptrsv_row_kernel(
    int nb, const int4 *__restrict__ info, const int *__restrict__ cx, const double *__restrict__ S,
    const double *__restrict__ dinv, const double *rhs, double *out, double *sentinel_fill, Counters *cnt,
    const int *__restrict__ skip) {
    if (skip && *skip) return;
    __shared__ int s_t;
    extern __shared__ unsigned long long row_x[];  // kRowsPerWG x 4 published values
    unsigned int *ticket = UPPER ? &cnt->ticket_u : &cnt->ticket_l;
    unsigned int *donec = UPPER ? &cnt->done_u : &cnt->done_l;
    const int limit = cnt->spin_limit;
    if (threadIdx.x == 0) s_t = (int)atomicAdd(ticket, 1u);
    row_x[threadIdx.x] = kSentinel;
    __syncthreads();
    const int q = s_t * kRowsPerWG + (int)(threadIdx.x >> 2);
    const int m = threadIdx.x & 3;
    const bool active = q < nb;
    int i = 0, pb = 0, pe = 0;
    double z = 0.0;
    if (active) {
        const int4 f = info[q];
        i = f.x, pb = f.y, pe = f.z;
        const double *D = dinv + (size_t)i * 16 + m * 4;
        const double *b = rhs + (size_t)i * 4;
        z = D[0] * b[0] + D[1] * b[1] + D[2] * b[2] + D[3] * b[3];
    }
    // batch registers: where each dependency's value is polled (its entry of the
    // solution vector, or its LDS slot), row m of its block, polled x_k[m]
    unsigned long long *src[kRowBatch];
    double s0[kRowBatch], s1[kRowBatch], s2[kRowBatch], s3[kRowBatch];
    unsigned long long u[kRowBatch];
    unsigned miss = 0;
    int p0 = pb;
    // One poll of every input in `want`: all loads are issued before any result is
    // compared, so a round costs one round trip however many inputs are missing
    // (a compare next to each load makes the compiler wait for every load in turn).
    // The loads are flat loads of addresses in LDS or in the solution vector, so
    // the round is straight-line code (separate LDS and global loads under
    // divergent branches made the compiler wait between them).  Returns whether
    // an input arrived.
    auto poll_round = [&](unsigned want) {
#pragma unroll
        for (int j = 0; j < kRowBatch; ++j)
            if (want & (1u << j)) u[j] = __hip_atomic_load(src[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned before = miss;
#pragma unroll
        for (int j = 0; j < kRowBatch; ++j)
            if ((want & (1u << j)) && u[j] != kSentinel) miss &= ~(1u << j);
        return miss != before;
    };
    auto load_batch = [&]() {
        const int n = min(kRowBatch, pe - p0);
        int kk[kRowBatch];
#pragma unroll
        for (int j = 0; j < kRowBatch; ++j) {
            kk[j] = cx[p0 + min(j, n - 1)];  // unconditional: no register copy waits for it
            if (j < n) {
                const double *B = S + (size_t)(p0 + j) * 16 + m * 4;
                s0[j] = B[0], s1[j] = B[1], s2[j] = B[2], s3[j] = B[3];
            }
        }
#pragma unroll
        for (int j = 0; j < kRowBatch; ++j)
            if (j < n)
                src[j] = kk[j] >= 0 ? (unsigned long long *)(out + (size_t)kk[j] * 4 + m) : &row_x[(~kk[j]) * 4 + m];
        // every address before the first poll: a flat load in flight makes the
        // compiler wait on both counters before the next address computation
        __builtin_amdgcn_sched_barrier(0);
        miss = (1u << n) - 1u;
        poll_round(miss);
    };
    if (p0 < pe) load_batch();
    bool done_row = !active;
    bool timed_out = false;
    int stall = 0;
    while (true) {
        bool progressed = false;
        const bool ready = quad_all(miss == 0);  // all lanes take part in the DPP moves
        if (!done_row && ready) {
            // every input of this batch is published: x_k arrives by quad broadcast
#pragma unroll
            for (int j = 0; j < kRowBatch; ++j)
                if (p0 + j < pe) {
                    const double xm = __longlong_as_double((long long)u[j]);
                    const double x0 = quad_bcast<0>(xm), x1 = quad_bcast<1>(xm);
                    const double x2 = quad_bcast<2>(xm), x3 = quad_bcast<3>(xm);
                    z -= s0[j] * x0 + s1[j] * x1 + s2[j] * x2 + s3[j] * x3;
                }
            p0 += kRowBatch;
            if (p0 < pe) {
                load_batch();
            } else {
                __hip_atomic_store(&row_x[threadIdx.x], (unsigned long long)__double_as_longlong(z), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
                publish(out + (size_t)i * 4 + m, z);
                if (sentinel_fill) sentinel_fill[(size_t)i * 4 + m] = __longlong_as_double((long long)kSentinel);
                done_row = true;
            }
            progressed = true;
        }
        if (__all(done_row)) break;
        // one poll of every missing input of the current batch (polling only the
        // input expected last, and the others once it is in, cost a round trip
        // more per level than it saved in fabric load)
        if (!done_row && miss && poll_round(miss)) progressed = true;
        stall = __any(progressed) ? 0 : stall + 1;
        if (stall > limit) {
            // give up on the missing inputs: they enter the row as NaN
            if (miss) timed_out = true;
            miss = 0;
            stall = 0;
        }
    }
    if (__any(timed_out) && threadIdx.x % 64 == 0) atomicOr(&cnt->err, 1);
    if (threadIdx.x == 0) {
        if (atomicAdd(donec, 1u) == (unsigned)gridDim.x - 1) {
            atomicExch(ticket, 0u);
            atomicExch(donec, 0u);
        }
    }
}

// ---------------------------------------------- BSR, bs = 4, unit schedule
// The row schedule's order cut into units of 64 rows (one workgroup each) grown
// so that most dependencies fall inside a unit.  A workgroup has four compute
// waves (one quad per row, as above) and one import wave:
//   * the import wave polls the solution vector for the unit's imports (the
//     distinct rows of other units its rows depend on) and copies each value into
//     an LDS slot as soon as it is published;
//   * the compute waves only ever poll LDS -- the unit's own rows' slots and the
//     import slots -- so a pass of a compute wave costs an LDS round trip, not
//     the L2 round trip of an agent-scope poll, and a row whose inputs are inside
//     the unit follows its producer within a few hundred nanoseconds.
// cxu[p] = ~slot for every dependency: slot < 64 is a row of the unit, 64 + k its
// k-th import.  Imports of unit w are imp_col[imp_ptr[w] .. imp_ptr[w+1]).
// kUnitBatch dependencies are held per pass, all polled every pass without a
// branch; the plan takes the smallest of 2, 4, 8 that covers the longest row
// (96^3 7-point, 3 dependencies: 840 us for L+U with 4 against 883 with 8;
// 64^3 27-point, 13: 2.78 ms with 8 against 3.51 with 4).
constexpr int kUnitWG = kRowWG + 64;
constexpr int kUnitImpMax = 512;                 // imports per unit (2048 LDS words)
// residency cap: two workgroups per CU (96^3 7-point L+U: 827 us; three 838, four 838)
constexpr int kUnitLds = 64 * 1024;
constexpr int kUnitItemsPerLane = kUnitImpMax * 4 / 64;

// The import wave of a unit: lane l owns items e = l + 64 r (import e / 4,
// component e % 4), polls the solution vector for each item until it is published
// and copies it into slot[e] (LDS).  All polls of a group of eight items are
// issued before any is compared.  Gives up after `limit` passes without progress
// (the compute waves then time out on the missing slots); returns whether it did.
__device__ __forceinline__ bool import_wave(int l, int nit, const int *__restrict__ imp, const double *out,
                                            unsigned long long *slot, int limit) {
    // word of the solution vector behind each item; the loads are issued
    // together (clamped, no default value to merge, so no wait between them)
    int at[kUnitItemsPerLane];
    unsigned pend = 0;
    if (nit > 0) {
#pragma unroll
        for (int r = 0; r < kUnitItemsPerLane; ++r) {
            const int e = min(l + 64 * r, nit - 1);
            at[r] = imp[e >> 2] * 4 + (e & 3);
            if (l + 64 * r < nit) pend |= 1u << r;
        }
    }
    bool timed_out = false;
    int stall = 0;
    constexpr int kG = 8;  // items polled before any is compared
    while (__any(pend != 0)) {
        bool got = false;
#pragma unroll
        for (int g = 0; g < kUnitItemsPerLane / kG; ++g) {
            const unsigned pm = (pend >> (kG * g)) & ((1u << kG) - 1u);
            if (!__any(pm != 0)) continue;
            unsigned long long v[kG];
#pragma unroll
            for (int j = 0; j < kG; ++j)
                if (pm & (1u << j))
                    v[j] = __hip_atomic_load((unsigned long long *)(out + at[kG * g + j]), __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
            for (int j = 0; j < kG; ++j)
                if ((pm & (1u << j)) && v[j] != kSentinel) {
                    __hip_atomic_store(&slot[l + 64 * (kG * g + j)], v[j], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                    pend &= ~(1u << (kG * g + j));
                    got = true;
                }
        }
        stall = __any(got) ? 0 : stall + 1;
        if (stall > limit) {
            // give up: the compute waves time out on the missing slots
            if (pend) timed_out = true;
            pend = 0;
        }
    }
    return timed_out;
}

template <bool UPPER, int kUnitBatch>
__global__ __launch_bounds__(kUnitWG) void bsr4_sptrsv_unit_kernel(
    int nb, const int4 *__restrict__ info, const int *__restrict__ cxu, const int *__restrict__ imp_ptr,
    const int *__restrict__ imp_col, const double *__restrict__ S, const double *__restrict__ dinv,
    const double *rhs, double *out, double *sentinel_fill, Counters *cnt, const int *__restrict__ skip) {
    if (skip && *skip) return;
    __shared__ int s_t;
    extern __shared__ unsigned long long unit_x[];  // (64 + imports) x 4 values
    unsigned int *ticket = UPPER ? &cnt->ticket_u : &cnt->ticket_l;
    unsigned int *donec = UPPER ? &cnt->done_u : &cnt->done_l;
    const int limit = cnt->spin_limit;
    if (threadIdx.x == 0) s_t = (int)atomicAdd(ticket, 1u);
    __syncthreads();
    const int w = s_t;
    const int ib = imp_ptr[w], nit = (imp_ptr[w + 1] - ib) * 4;
    for (int t = threadIdx.x; t < kRowWG + nit; t += kUnitWG) unit_x[t] = kSentinel;
    __syncthreads();
    bool timed_out = false;
    if (threadIdx.x >= kRowWG) {
        timed_out = import_wave(threadIdx.x - kRowWG, nit, imp_col + ib, out, unit_x + kRowWG, limit);
    } else {
        const int q = w * kRowsPerWG + (int)(threadIdx.x >> 2);
        const int m = threadIdx.x & 3;
        const bool active = q < nb;
        int i = 0, pb = 0, pe = 0;
        double z = 0.0;
        if (active) {
            const int4 f = info[q];
            i = f.x, pb = f.y, pe = f.z;
            const double *D = dinv + (size_t)i * 16 + m * 4;
            const double *b = rhs + (size_t)i * 4;
            z = D[0] * b[0] + D[1] * b[1] + D[2] * b[2] + D[3] * b[3];
        }
        int sl[kUnitBatch];  // LDS word of each dependency's component m
        double s0[kUnitBatch], s1[kUnitBatch], s2[kUnitBatch], s3[kUnitBatch];
        unsigned long long u[kUnitBatch];
        unsigned miss = 0;
        int p0 = pb;
        // every slot of the batch is read on every pass (clamped slots for the
        // unused ones): no branch per input, and an arrived value reads the same
        auto poll_round = [&]() {
#pragma unroll
            for (int j = 0; j < kUnitBatch; ++j)
                u[j] = __hip_atomic_load(&unit_x[sl[j]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            unsigned got = 0;
#pragma unroll
            for (int j = 0; j < kUnitBatch; ++j) got |= (u[j] != kSentinel ? 1u : 0u) << j;
            const unsigned before = miss;
            miss &= ~got;
            return miss != before;
        };
        auto load_batch = [&]() {
            const int n = min(kUnitBatch, pe - p0);
#pragma unroll
            for (int j = 0; j < kUnitBatch; ++j) {
                sl[j] = (~cxu[p0 + min(j, n - 1)]) * 4 + m;
                if (j < n) {
                    const double *B = S + (size_t)(p0 + j) * 16 + m * 4;
                    s0[j] = B[0], s1[j] = B[1], s2[j] = B[2], s3[j] = B[3];
                }
            }
            miss = (1u << n) - 1u;
            poll_round();
        };
        if (p0 < pe) load_batch();
        bool done_row = !active;
        int stall = 0;
        long long t_stall = 0;
        while (true) {
            bool progressed = false;
            const bool ready = quad_all(miss == 0);
            if (!done_row && ready) {
#pragma unroll
                for (int j = 0; j < kUnitBatch; ++j)
                    if (p0 + j < pe) {
                        const double xm = __longlong_as_double((long long)u[j]);
                        const double x0 = quad_bcast<0>(xm), x1 = quad_bcast<1>(xm);
                        const double x2 = quad_bcast<2>(xm), x3 = quad_bcast<3>(xm);
                        z -= s0[j] * x0 + s1[j] * x1 + s2[j] * x2 + s3[j] * x3;
                    }
                p0 += kUnitBatch;
                if (p0 < pe) {
                    load_batch();
                } else {
                    __hip_atomic_store(&unit_x[threadIdx.x], (unsigned long long)__double_as_longlong(z),
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    publish(out + (size_t)i * 4 + m, z);
                    if (sentinel_fill) sentinel_fill[(size_t)i * 4 + m] = __longlong_as_double((long long)kSentinel);
                    done_row = true;
                }
                progressed = true;
            }
            if (__all(done_row)) break;
            if (!done_row && miss && poll_round()) progressed = true;
            // LDS passes take tens of nanoseconds: the wait is bounded by the clock
            // (100 MHz), spin_limit microseconds without progress
            if (__any(progressed)) {
                stall = 0;
            } else {
                if (stall == 0) t_stall = __builtin_amdgcn_s_memrealtime();
                if ((++stall & 255) == 0 && __builtin_amdgcn_s_memrealtime() - t_stall > (long long)limit * 100) {
                    if (miss) timed_out = true;
                    miss = 0;
                    stall = 0;
                }
            }
        }
    }
    if (__any(timed_out) && threadIdx.x % 64 == 0) atomicOr(&cnt->err, 1);
    if (threadIdx.x == 0) {
        if (atomicAdd(donec, 1u) == (unsigned)gridDim.x - 1) {
            atomicExch(ticket, 0u);
            atomicExch(donec, 0u);
        }
    }
}

// ------------------------------------------------ level-synchronous units
// Units of kLvUnit = 512 rows, grown by the unit order below with that size (on a
// 3-D 7-point grid in natural order the greedy growth cuts the grid into 8x8x8
// blocks: 88 % of the dependencies inside a unit, 22 local levels, at most 48 rows
// per level and 192 imports per unit).  Inside a unit the rows are sorted by their
// local level (1 + the deepest local level of their dependencies in the unit), and
// the unit's eight compute waves step through the levels together: a row's inputs
// from its own unit were written to LDS at a lower level, so after the level
// barrier they are read without polling; inputs from other units arrive through
// the import wave as in the unit kernel.  A level costs one barrier and one row
// computation, where the unit kernel's per-row LDS polling cost about 0.8 us per
// hop, and a dependency path of the 96^3 factor crosses 33 units instead of 69.
//
// Compute quad q (of 128) owns the rows at slots q + 128 k, k < 4, and loads their
// inputs (block row m of each S_ik) into registers at the unit's start, and the
// unit's values inv(D_i) rhs_i are formed in LDS then, so no global load lies on a
// level's path.  A wave's quads hold 16 consecutive slots of each k; the analysis
// checks that every wave meets its rows in increasing level order (the highest
// level of its rows k - 1 below the lowest of its rows k), so one level loop per k
// holds one row's code, and that no row has more than B (at most 3) dependencies,
// else the plan uses the unit kernel.  The level barrier of the
// compute waves is an LDS counter (the import wave keeps polling; s_barrier would
// tie it to the levels), preceded by a wait for the wave's LDS stores only: the
// global stores that publish values are not waited for.
constexpr int kLvUnit = 512;
constexpr int kLvQuads = 128;
constexpr int kLvRows = kLvUnit / kLvQuads;  // rows per quad
constexpr int kLvCompute = kLvQuads * 4;
constexpr int kLvWG = kLvCompute + 64;
constexpr int kLvWaves = kLvCompute / 64;

template <bool UPPER, int B>
__global__ __launch_bounds__(kLvWG) void bsr4_sptrsv_level_kernel(
    int nb, const int4 *__restrict__ info, const int *__restrict__ cxu, const int *__restrict__ imp_ptr,
    const int *__restrict__ imp_col, const double *__restrict__ S, const double *__restrict__ dinv,
    const double *rhs, double *out, double *sentinel_fill, Counters *cnt, const int *__restrict__ skip,
    int imp_cap) {
    if (skip && *skip) return;
    __shared__ int s_t;
    __shared__ unsigned int s_bar;
    // (kLvUnit + imp_cap) x 4 values: the unit's rows, then its imports; then the
    // unit's kLvUnit x 4 values of c = inv(D_i) rhs_i
    extern __shared__ unsigned long long lv_x[];
    unsigned int *ticket = UPPER ? &cnt->ticket_u : &cnt->ticket_l;
    unsigned int *donec = UPPER ? &cnt->done_u : &cnt->done_l;
    const int limit = cnt->spin_limit;
    if (threadIdx.x == 0) {
        s_t = (int)atomicAdd(ticket, 1u);
        s_bar = 0u;
    }
    __syncthreads();
    const int w = s_t;
    const int ib = imp_ptr[w], nit = (imp_ptr[w + 1] - ib) * 4;
    for (int t = threadIdx.x; t < nit; t += kLvWG) lv_x[kLvUnit * 4 + t] = kSentinel;
    const int zero_slot = (2 * kLvUnit + imp_cap) * 4;  // four zero words after the c values
    if (threadIdx.x < 4) lv_x[zero_slot + threadIdx.x] = 0ull;
    __syncthreads();
    bool timed_out = false;
    if (threadIdx.x >= kLvCompute) {
        timed_out = import_wave(threadIdx.x - kLvCompute, nit, imp_col + ib, out, lv_x + kLvUnit * 4, limit);
    } else {
        const int quad = threadIdx.x >> 2, m = threadIdx.x & 3, wv = threadIdx.x >> 6;
        const int base = w * kLvUnit;
        const int nrows = min(kLvUnit, nb - base);
        // The unit's inputs are loaded in two rounds of unconditional loads (clamped
        // positions and blocks; a missing row reads valid memory and is ignored):
        // first the position records, then every factor block, inverse diagonal
        // block and right-hand side they point at.  Loads that wait on each other
        // row by row took 14.6 us per unit, most of its time in the workgroup slot.
        // Compute thread t forms c = inv(D_i) rhs_i of the unit's row at slot t in LDS;
        // quad q keeps row m of each S_ik of its rows at slots q + 128 k, level
        // (INT_MAX: none), row and the LDS word of each dependency's component m.
        double *cbuf = (double *)(lv_x + (kLvUnit + imp_cap) * 4);
        // unused inputs of a row read four zeros at zero_slot, with zero blocks in s,
        // and subtract an exact zero
        const int nl = info[base + nrows - 1].w + 1;
        const int4 ft = info[base + min((int)threadIdx.x, nrows - 1)];
        int4 f[kLvRows];
        int lvl[kLvRows], irow[kLvRows], nd[kLvRows], hi[kLvRows];
#pragma unroll
        for (int k = 0; k < kLvRows; ++k) {
            f[k] = info[base + min(quad + kLvQuads * k, nrows - 1)];
            // highest level among the wave's rows k (its slots are consecutive and
            // sorted by level), -1 when the wave has none
            const int first = kLvQuads * k + 16 * wv;
            hi[k] = __builtin_amdgcn_readfirstlane(first < nrows ? info[base + min(first + 15, nrows - 1)].w : -1);
        }
        int sl[kLvRows][B];  // LDS word of each input's component 0 (packed in pairs below)
        double s[kLvRows][B][4];
#pragma unroll
        for (int k = 0; k < kLvRows; ++k) {
            const bool has = quad + kLvQuads * k < nrows;
            lvl[k] = has ? f[k].w : INT_MAX;
            irow[k] = f[k].x;
            nd[k] = has ? f[k].z - f[k].y : 0;
#pragma unroll
            for (int j = 0; j < B; ++j) {
                const int pj = j < nd[k] ? f[k].y + j : 0;
                sl[k][j] = cxu[pj];
                const double *Bk = S + (size_t)pj * 16 + m * 4;
                s[k][j][0] = Bk[0], s[k][j][1] = Bk[1], s[k][j][2] = Bk[2], s[k][j][3] = Bk[3];
            }
            // an unused input gets a zero block (whatever block 0 of S holds)
#pragma unroll
            for (int j = 0; j < B; ++j)
                if (j >= nd[k]) s[k][j][0] = s[k][j][1] = s[k][j][2] = s[k][j][3] = 0.0;
        }
        {
            const double *D = dinv + (size_t)ft.x * 16;
            const double *b = rhs + (size_t)ft.x * 4;
            double Dv[16], bv[4];
#pragma unroll
            for (int e = 0; e < 16; ++e) Dv[e] = D[e];
#pragma unroll
            for (int e = 0; e < 4; ++e) bv[e] = b[e];
            // every slot is written (slots past the unit's rows are never read), so
            // the wait for these loads lies on every path into the level loop
#pragma unroll
            for (int r = 0; r < 4; ++r)
                cbuf[threadIdx.x * 4 + r] =
                    Dv[r * 4] * bv[0] + Dv[r * 4 + 1] * bv[1] + Dv[r * 4 + 2] * bv[2] + Dv[r * 4 + 3] * bv[3];
        }
#pragma unroll
        for (int k = 0; k < kLvRows; ++k)
#pragma unroll
            for (int j = 0; j < B; ++j) sl[k][j] = j < nd[k] ? (~sl[k][j]) * 4 : zero_slot;  // component 0
        unsigned slp[kLvRows][(B + 1) / 2];
#pragma unroll
        for (int k = 0; k < kLvRows; ++k)
#pragma unroll
            for (int j = 0; j < B; j += 2)
                slp[k][j / 2] = (unsigned)sl[k][j] | (j + 1 < B ? (unsigned)sl[k][j + 1] << 16 : 0u);
        // every load of the set-up is complete here: the compiler's waits for them
        // would otherwise fall inside the level loop, where each also waits for the
        // stores that publish the previous levels
        __builtin_amdgcn_s_waitcnt(0);
        // one row of this quad: read its inputs (own-unit inputs are final after the
        // previous level's barrier; imports are polled in LDS), compute, publish
        auto row = [&](int k) {
            // component m of each input (unused inputs read a zero), then the quad
            // exchanges components by DPP; an import still in flight reads the
            // sentinel and is read again
            unsigned long long u[B];
            unsigned miss = 0;
#pragma unroll
            for (int j = 0; j < B; ++j) {
                u[j] = __hip_atomic_load(&lv_x[((slp[k][j / 2] >> (16 * (j & 1))) & 0xffffu) + m], __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_WORKGROUP);
                if (u[j] == kSentinel) miss |= 1u << j;
            }
            int stall = 0;
            long long t_stall = 0;
            while (!quad_all(miss == 0)) {
#pragma unroll
                for (int j = 0; j < B; ++j)
                    if (miss & (1u << j)) {
                        u[j] = __hip_atomic_load(&lv_x[((slp[k][j / 2] >> (16 * (j & 1))) & 0xffffu) + m],
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        if (u[j] != kSentinel) miss &= ~(1u << j);
                    }
                // bounded by the clock (100 MHz): spin_limit microseconds
                if (stall == 0) t_stall = __builtin_amdgcn_s_memrealtime();
                if ((++stall & 255) == 0 && __builtin_amdgcn_s_memrealtime() - t_stall > (long long)limit * 100) {
                    if (miss) timed_out = true;
                    miss = 0;
                }
            }
            double z = cbuf[(quad + kLvQuads * k) * 4 + m];
#pragma unroll
            for (int j = 0; j < B; ++j) {
                const double xm = __longlong_as_double((long long)u[j]);
                const double x0 = quad_bcast<0>(xm), x1 = quad_bcast<1>(xm);
                const double x2 = quad_bcast<2>(xm), x3 = quad_bcast<3>(xm);
                z -= s[k][j][0] * x0 + s[k][j][1] * x1 + s[k][j][2] * x2 + s[k][j][3] * x3;
            }
            const int slot = quad + kLvQuads * k;
            __hip_atomic_store(&lv_x[slot * 4 + m], (unsigned long long)__double_as_longlong(z), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
            publish(out + (size_t)irow[k] * 4 + m, z);
            if (sentinel_fill) sentinel_fill[(size_t)irow[k] * 4 + m] = __longlong_as_double((long long)kSentinel);
        };
        // the compute waves' barrier after level l: this wave's LDS stores first
        // (not its global stores), then one arrival per wave and a bounded wait
        auto level_barrier = [&](int l) {
            __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
            if ((threadIdx.x & 63) == 0) __hip_atomic_fetch_add(&s_bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            const unsigned target = (unsigned)kLvWaves * (unsigned)(l + 2);  // round 0: the set-up
            int stall = 0;
            long long t0 = 0;
            while (__hip_atomic_load(&s_bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < target) {
                if (stall == 0) t0 = __builtin_amdgcn_s_memrealtime();
                if ((++stall & 255) == 0 && __builtin_amdgcn_s_memrealtime() - t0 > (long long)limit * 100) {
                    timed_out = true;
                    break;
                }
            }
            __asm__ volatile("" ::: "memory");
        };
        level_barrier(-1);  // the c values of every row are in LDS
        // the wave meets its rows k = 0, 1, 2, 3 in increasing level order (checked
        // by the analysis), so one level loop per k holds one row's code
        int l = 0;
#pragma unroll
        for (int k = 0; k < kLvRows; ++k) {
            for (; l <= hi[k]; ++l) {
                if (lvl[k] == l) row(k);
                level_barrier(l);
            }
        }
        for (; l < nl; ++l) level_barrier(l);
    }
    if (__any(timed_out) && threadIdx.x % 64 == 0) atomicOr(&cnt->err, 1);
    if (threadIdx.x == 0) {
        if (atomicAdd(donec, 1u) == (unsigned)gridDim.x - 1) {
            atomicExch(ticket, 0u);
            atomicExch(donec, 0u);
        }
    }
}

// S_p = inv(D_i) T_p for every off-diagonal block p of row i (T = L below the
// diagonal, U above).  One wave per block row, one lane per element of the row's
// blocks, so the block reads and writes are whole 128-byte lines (one thread per
// row, walking its blocks, took 873 us on the 96^3 7-pt factor).
constexpr int kScaleBlock = 256;
__global__ __launch_bounds__(kScaleBlock) void scale_offdiag_kernel(int nb, const int *__restrict__ rp,
                                                                    const int *__restrict__ diagpos,
                                                                    const double *__restrict__ LU,
                                                                    const double *__restrict__ dinvL,
                                                                    const double *__restrict__ dinvU,
                                                                    double *__restrict__ S) {
    const int i = blockIdx.x * (kScaleBlock / 64) + (int)(threadIdx.x >> 6);
    if (i >= nb) return;
    const int lane = threadIdx.x & 63;
    const int dp = diagpos[i], p0 = rp[i], len = rp[i + 1] - p0;
    for (int t = lane; t < len * 16; t += 64) {
        const int p = p0 + (t >> 4), e = t & 15, r = e >> 2, c = e & 3;
        if (p == dp) continue;
        const double *D = (p < dp ? dinvL : dinvU) + (size_t)i * 16 + r * 4;
        const double *T = LU + (size_t)p * 16 + c;
        S[(size_t)p * 16 + e] = D[0] * T[0 * 4] + D[1] * T[1 * 4] + D[2] * T[2 * 4] + D[3] * T[3 * 4];
    }
}

// ------------------------------------------------------------- CSR, scalar
// cape semantics: L has a unit diagonal that is skipped (cape_sptrsv.h:36-47);
// U divides by the diagonal entry (cape_sptrsv.h:95-96).  The wave's 64 lanes
// split the entries of one row; the row sum is a wave reduction.
template <bool UPPER>
__global__ __launch_bounds__(64) void csr_sptrsv_kernel(int m, int chunk, int nchunks, const int *__restrict__ rp,
                                                        const int *__restrict__ ci, const double *__restrict__ v,
                                                        const double *rhs, double *out,
                                                        double *sentinel_fill, Counters *cnt) {
    __shared__ double xs[kCsrMaxChunk];
    __shared__ int s_c;
    const int lane = threadIdx.x;
    unsigned int *ticket = UPPER ? &cnt->ticket_u : &cnt->ticket_l;
    unsigned int *done = UPPER ? &cnt->done_u : &cnt->done_l;
    const int limit = cnt->spin_limit;
    if (lane == 0) s_c = (int)atomicAdd(ticket, 1u);
    __syncthreads();
    const int t = s_c;
    const int c = UPPER ? nchunks - 1 - t : t;
    const int lo = c * chunk, hi = min(lo + chunk, m);
    bool timed_out = false;
    for (int step = 0; step < hi - lo; ++step) {
        const int i = UPPER ? hi - 1 - step : lo + step;
        double acc = 0.0, diag = 0.0;
        int has_diag = 0;
        for (int p = rp[i] + lane; p < rp[i + 1]; p += 64) {
            const int k = ci[p];
            if (k == i) {
                diag = v[p];
                has_diag = 1;
                continue;
            }
            if (UPPER ? (k < i) : (k > i)) continue;
            double xk;
            const bool local = UPPER ? (k < hi) : (k >= lo);
            if (local) xk = xs[k - lo];
            else xk = poll(out + k, limit, timed_out);
            acc = fma(v[p], xk, acc);
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off);
        double val = rhs[i] - acc;
        if (UPPER) {
            const unsigned long long bal = __ballot(has_diag);
            const int src = bal ? __ffsll((long long)bal) - 1 : 0;
            const double d = __shfl(diag, src);
            if (!bal) timed_out = true;  // no diagonal entry: unsolvable row
            val = val / d;
        }
        if (lane == 0) {
            xs[i - lo] = val;
            publish(out + i, val);
            if (sentinel_fill) sentinel_fill[i] = __longlong_as_double((long long)kSentinel);
        }
        __syncthreads();
    }
    if (timed_out) atomicOr(&cnt->err, 1);
    if (lane == 0) {
        if (atomicAdd(done, 1u) == (unsigned)nchunks - 1) {
            atomicExch(ticket, 0u);
            atomicExch(done, 0u);
        }
    }
}

__global__ void fill_sentinel_kernel(double *p, size_t n, const int *__restrict__ skip) {
    if (skip && *skip) return;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        p[k] = __longlong_as_double((long long)kSentinel);
}

static unsigned fill_blocks(size_t n) {
    size_t g = (n + 255) / 256;
    return (unsigned)(g > 8192 ? 8192 : (g ? g : 1));
}

}  // namespace isai

using namespace isai;

struct isai_sptrsv_plan {
    int nb = 0, nnzb = 0, bs = 0, chunk = 0;
    bool rows = false;   // row schedule (one quad per block row) instead of chunks
    bool units = false;        // row schedule in unit order (else level order)
    bool unit_kernel = false;  // units solved by the import-wave kernel (else the row kernel)
    bool level_kernel = false; // level-synchronous units of kLvUnit rows (bsr4_sptrsv_level_kernel)
    int lv_batch = 3;          // level kernel: dependencies held per row (2 or 3)
    int4 *d_info[2] = {nullptr, nullptr};  // row schedule: (row, first, end, 1) per position
    int *d_cx = nullptr;  // row kernel: column of each dependency block, ~slot inside its workgroup;
                          // unit kernel: ~slot of each dependency (unit row or import)
    int *d_imp_ptr[2] = {nullptr, nullptr}, *d_imp_col[2] = {nullptr, nullptr};  // unit kernel imports
    int row_lds = 0;     // dynamic LDS per workgroup of the row or unit kernel
    int unit_batch = 4;  // unit kernel: dependencies held per pass (2, 4 or 8)
    int poll_min = 64;  // smallest poll window of the adaptive poll (64 = fixed 64-item window)
    int nchunks[2] = {0, 0};
    int maxr[2] = {0, 0}, maxrem[2] = {0, 0}, maxloc[2] = {0, 0};
    const int *d_row = nullptr, *d_col = nullptr;
    int *d_diag = nullptr;
    double *d_dinvL = nullptr, *d_dinvU = nullptr, *d_y = nullptr, *d_S = nullptr;
    int4 *d_chunks[2] = {nullptr, nullptr};
    int *d_rem_ptr[2] = {nullptr, nullptr}, *d_rem_items[2] = {nullptr, nullptr};
    int *d_loc_ptr[2] = {nullptr, nullptr}, *d_loc_items[2] = {nullptr, nullptr};
    Counters *d_cnt = nullptr;
};

namespace isai {

template <int B>
static void unit_launch(isai_sptrsv_plan_t P, const double *d_b, double *d_x, hipStream_t st, const int *skip) {
    const unsigned g = (unsigned)((P->nb + kRowsPerWG - 1) / kRowsPerWG);
    bsr4_sptrsv_unit_kernel<false, B><<<g, kUnitWG, P->row_lds, st>>>(P->nb, P->d_info[0], P->d_cx, P->d_imp_ptr[0],
                                                                     P->d_imp_col[0], P->d_S, P->d_dinvL, d_b,
                                                                     P->d_y, d_x, P->d_cnt, skip);
    bsr4_sptrsv_unit_kernel<true, B><<<g, kUnitWG, P->row_lds, st>>>(P->nb, P->d_info[1], P->d_cx, P->d_imp_ptr[1],
                                                                    P->d_imp_col[1], P->d_S, P->d_dinvU, P->d_y,
                                                                    d_x, nullptr, P->d_cnt, skip);
}

template <int B>
static void level_launch(isai_sptrsv_plan_t P, const double *d_b, double *d_x, hipStream_t st, const int *skip) {
    const unsigned g = (unsigned)((P->nb + kLvUnit - 1) / kLvUnit);
    const int cap = P->row_lds / 32 - 2 * kLvUnit - 1;  // import slots of the LDS layout
    bsr4_sptrsv_level_kernel<false, B><<<g, kLvWG, P->row_lds, st>>>(P->nb, P->d_info[0], P->d_cx,
                                                                     P->d_imp_ptr[0], P->d_imp_col[0], P->d_S,
                                                                     P->d_dinvL, d_b, P->d_y, d_x, P->d_cnt, skip,
                                                                     cap);
    bsr4_sptrsv_level_kernel<true, B><<<g, kLvWG, P->row_lds, st>>>(P->nb, P->d_info[1], P->d_cx, P->d_imp_ptr[1],
                                                                    P->d_imp_col[1], P->d_S, P->d_dinvU, P->d_y,
                                                                    d_x, nullptr, P->d_cnt, skip, cap);
}

template <int B>
static hipError_t level_attr(int lds) {
    hipError_t e = hipFuncSetAttribute((const void *)bsr4_sptrsv_level_kernel<false, B>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)bsr4_sptrsv_level_kernel<true, B>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    return e;
}

template <int B>
static hipError_t unit_attr(int lds) {
    hipError_t e = hipFuncSetAttribute((const void *)bsr4_sptrsv_unit_kernel<false, B>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)bsr4_sptrsv_unit_kernel<true, B>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    return e;
}

int sptrsv_solve_internal(isai_sptrsv_plan_t P, const double *d_LU, const double *d_b, double *d_x,
                          hipStream_t st, const int *skip) {
    (void)d_LU;  // the solve reads the pre-scaled copy made from these values
    const size_t n = (size_t)P->nb * 4;
    fill_sentinel_kernel<<<fill_blocks(n), 256, 0, st>>>(P->d_y, n, skip);
    if (P->level_kernel) {
        if (P->lv_batch == 2)
            level_launch<2>(P, d_b, d_x, st, skip);
        else
            level_launch<3>(P, d_b, d_x, st, skip);
        ISAI_LAUNCH_CHECK();
        return ISAI_OK;
    }
    if (P->unit_kernel) {
        if (P->unit_batch == 2)
            unit_launch<2>(P, d_b, d_x, st, skip);
        else if (P->unit_batch == 4)
            unit_launch<4>(P, d_b, d_x, st, skip);
        else
            unit_launch<8>(P, d_b, d_x, st, skip);
        ISAI_LAUNCH_CHECK();
        return ISAI_OK;
    }
    if (P->rows) {
        const unsigned g = (unsigned)((P->nb + kRowsPerWG - 1) / kRowsPerWG);
        bsr4_sptrsv_row_kernel<false><<<g, kRowWG, P->row_lds, st>>>(P->nb, P->d_info[0], P->d_cx, P->d_S,
                                                                  P->d_dinvL, d_b, P->d_y, d_x, P->d_cnt, skip);
        bsr4_sptrsv_row_kernel<true><<<g, kRowWG, P->row_lds, st>>>(P->nb, P->d_info[1], P->d_cx, P->d_S,
                                                                 P->d_dinvU, P->d_y, d_x, nullptr, P->d_cnt, skip);
        ISAI_LAUNCH_CHECK();
        return ISAI_OK;
    }
    bsr4_sptrsv_pipe_kernel<false><<<P->nchunks[0], 64, pipe_lds_bytes(P->maxr[0], P->maxrem[0], P->maxloc[0]), st>>>(
        P->nb, P->d_chunks[0], P->d_col, P->d_S, P->d_dinvL, P->d_rem_ptr[0], P->d_rem_items[0], P->d_loc_ptr[0],
        P->d_loc_items[0], d_b, P->d_y, d_x, P->d_cnt, skip, P->maxr[0], P->maxrem[0], P->maxloc[0], P->poll_min);
    bsr4_sptrsv_pipe_kernel<true><<<P->nchunks[1], 64, pipe_lds_bytes(P->maxr[1], P->maxrem[1], P->maxloc[1]), st>>>(
        P->nb, P->d_chunks[1], P->d_col, P->d_S, P->d_dinvU, P->d_rem_ptr[1], P->d_rem_items[1], P->d_loc_ptr[1],
        P->d_loc_items[1], P->d_y, d_x, nullptr, P->d_cnt, skip, P->maxr[1], P->maxrem[1], P->maxloc[1], P->poll_min);
    ISAI_LAUNCH_CHECK();
    return ISAI_OK;
}

// Chunk tables of one triangle (upper = false: L, ascending rows; true: U,
// descending rows), in processing positions q.
static int build_tables(int nb, const std::vector<int> &rp, const std::vector<int> &ci,
                        const std::vector<int> &diag, bool upper, int cap_rows, std::vector<int4> &chunks,
                        std::vector<int> &rem_ptr, std::vector<int> &rem_items, std::vector<int> &loc_ptr,
                        std::vector<int> &loc_items) {
    auto row_of = [&](int q) { return upper ? nb - 1 - q : q; };
    auto pos_of = [&](int i) { return upper ? nb - 1 - i : i; };
    auto tri = [&](int i, int &pb, int &pe) {
        pb = upper ? diag[i] + 1 : rp[i];
        pe = upper ? rp[i + 1] : diag[i];
    };
    // a row starts a chain when it does not depend on the row processed just before
    auto starts_chain = [&](int q) {
        if (q == 0) return true;
        const int i = row_of(q), prev = row_of(q - 1);
        int pb, pe;
        tri(i, pb, pe);
        for (int p = pb; p < pe; ++p)
            if (ci[p] == prev) return false;
        return true;
    };
    // remote / local item counts of row q inside a chunk starting at qlo
    auto counts = [&](int q, int qlo, int &nr, int &nl) {
        const int i = row_of(q);
        int pb, pe;
        tri(i, pb, pe);
        nr = nl = 0;
        for (int p = pb; p < pe; ++p) (pos_of(ci[p]) >= qlo ? nl : nr)++;
    };
    chunks.clear();
    rem_ptr.assign(nb + 1, 0);
    loc_ptr.assign(nb + 1, 0);
    rem_items.clear();
    loc_items.clear();
    int q = 0;
    while (q < nb) {
        const int qlo = q;
        int nrem = 0, nloc = 0;
        // grow the chunk row by row; stop at a chain start once the next chain
        // would not fit, or at the caps
        while (q < nb) {
            if (q > qlo && starts_chain(q)) {
                // does the whole next chain fit?
                int e = q + 1, r2 = nrem, l2 = nloc;
                bool fits = true;
                for (int t = q; t < nb && (t == q || !starts_chain(t)); ++t, e = t) {
                    int a, b;
                    counts(t, qlo, a, b);
                    r2 += a;
                    l2 += b;
                    if (t - qlo + 1 > cap_rows || r2 > kPipeRemote || l2 > kPipeLocal) { fits = false; break; }
                }
                (void)e;
                if (!fits) break;
            }
            int a, b;
            counts(q, qlo, a, b);
            if (a > 64) {
                set_error("sptrsv analysis: block row %d depends on %d blocks outside its chunk (limit 64)", row_of(q), a);
                return ISAI_ERR_UNSUPPORTED;
            }
            if (q > qlo && (q - qlo + 1 > cap_rows || nrem + a > kPipeRemote || nloc + b > kPipeLocal)) break;
            nrem += a;
            nloc += b;
            ++q;
        }
        chunks.push_back(make_int4(qlo, q, (int)rem_items.size(), (int)loc_items.size()));
        for (int t = qlo; t < q; ++t) {
            rem_ptr[t] = (int)rem_items.size();
            loc_ptr[t] = (int)loc_items.size();
            const int i = row_of(t);
            int pb, pe;
            tri(i, pb, pe);
            for (int p = pb; p < pe; ++p) (pos_of(ci[p]) >= qlo ? loc_items : rem_items).push_back(p);
        }
        // keep loc_ptr ordered: locals of row t come after remotes bookkeeping above
    }
    // the per-position pointers must list each row's own items: rebuild them in order
    rem_ptr.assign(nb + 1, 0);
    loc_ptr.assign(nb + 1, 0);
    {
        int r = 0, l = 0;
        size_t c = 0;
        for (int t = 0; t < nb; ++t) {
            while (c + 1 < chunks.size() && chunks[c + 1].x <= t) ++c;
            const int qlo = chunks[c].x;
            rem_ptr[t] = r;
            loc_ptr[t] = l;
            int a, b;
            counts(t, qlo, a, b);
            r += a;
            l += b;
        }
        rem_ptr[nb] = r;
        loc_ptr[nb] = l;
    }
    // Dispatch order: chunks sorted by their level in the chunk dependency graph
    // (1 + the deepest chunk any of their remote items lives in), rows order inside
    // a level.  Resident waves then hold the active wavefront instead of a block of
    // consecutive rows, and a chunk only waits on chunks with lower tickets.
    const int nch = (int)chunks.size();
    std::vector<int> chunk_of(nb), level(nch, 0);
    for (int c = 0; c < nch; ++c)
        for (int t = chunks[c].x; t < chunks[c].y; ++t) chunk_of[t] = c;
    int nlev = 0;
    for (int c = 0; c < nch; ++c) {
        const int rb = chunks[c].z, re = c + 1 < nch ? chunks[c + 1].z : (int)rem_items.size();
        int l = 0;
        for (int w = rb; w < re; ++w) l = std::max(l, level[chunk_of[pos_of(ci[rem_items[w]])]] + 1);
        level[c] = l;
        nlev = std::max(nlev, l + 1);
    }
    std::vector<int> order(nch);
    {
        std::vector<int> cnt(nlev + 1, 0);
        for (int c = 0; c < nch; ++c) cnt[level[c] + 1]++;
        for (int l = 0; l < nlev; ++l) cnt[l + 1] += cnt[l];
        for (int c = 0; c < nch; ++c) order[cnt[level[c]]++] = c;
    }
    std::vector<int4> out;
    out.reserve(2 * nch);
    for (int c : order) {
        const int re = c + 1 < nch ? chunks[c + 1].z : (int)rem_items.size();
        const int le = c + 1 < nch ? chunks[c + 1].w : (int)loc_items.size();
        out.push_back(chunks[c]);
        out.push_back(make_int4(re, le, level[c], 0));
    }
    chunks.swap(out);
    return ISAI_OK;
}

// Levels of one triangle: level(i) = 1 + the deepest level among the rows i
// depends on (0 without dependencies).  Returns the number of levels.
static int row_levels(int nb, const std::vector<int> &rp, const std::vector<int> &ci, const std::vector<int> &diag,
                      bool upper, std::vector<int> &lev) {
    lev.assign(nb, 0);
    int nlev = 0;
    for (int q = 0; q < nb; ++q) {
        const int i = upper ? nb - 1 - q : q;
        const int pb = upper ? diag[i] + 1 : rp[i], pe = upper ? rp[i + 1] : diag[i];
        int l = 0;
        for (int p = pb; p < pe; ++p) l = std::max(l, lev[ci[p]] + 1);
        lev[i] = l;
        nlev = std::max(nlev, l + 1);
    }
    return nlev;
}

// Level order of one triangle for the row schedule: rows by level, in
// processing-position order inside a level.
static void level_order(int nb, const std::vector<int> &lev, int nlev, bool upper, std::vector<int> &order) {
    std::vector<int> cnt(nlev + 1, 0);
    for (int i = 0; i < nb; ++i) cnt[lev[i] + 1]++;
    for (int l = 0; l < nlev; ++l) cnt[l + 1] += cnt[l];
    order.resize(nb);
    for (int q = 0; q < nb; ++q) {
        const int i = upper ? nb - 1 - q : q;
        order[cnt[lev[i]]++] = i;
    }
}

// Unit order of one triangle for the row schedule: the order is cut into units
// of usz consecutive rows (one workgroup each), and it is grown so that
// as many dependencies as possible fall inside a unit, where they travel through
// LDS.  Greedy: a unit starts at the shallowest row whose dependencies are all
// placed; it then repeatedly takes, among the rows that became ready while it
// grew, the one with the most dependencies inside the unit (the shallower level,
// then the processing position, on ties), and the shallowest ready row when
// none is left.  A row is placed only after all of its dependencies, so every
// dependency has a lower position (a lower or equal ticket).
//
// The walk moves along the level wavefront, which in row order spans the whole
// matrix (a 3-D stencil's level planes cut through every row range), so it first
// copies the triangle's dependency lists into level-major numbering, where the
// rows it touches at any time sit close together.  That numbering sorts rows by
// (level, position), the walk's own tie-break, so the ready pool is a two-level
// bitset over it (the shallowest ready row is its first set bit) and the rows that
// became ready in the current unit sit in a small heap keyed (-inside, number).
static void unit_order(int nb, const std::vector<int> &rp, const std::vector<int> &ci, const std::vector<int> &diag,
                       bool upper, const std::vector<int> &lev, int nlev, int usz, std::vector<int> &order) {
    // level-major numbering: perm[n] = row, inv[row] = n
    std::vector<int> cnt(nlev + 1, 0), perm(nb), inv(nb);
    for (int i = 0; i < nb; ++i) cnt[lev[i] + 1]++;
    for (int l = 0; l < nlev; ++l) cnt[l + 1] += cnt[l];
    for (int q = 0; q < nb; ++q) {
        const int i = upper ? nb - 1 - q : q;
        const int n = cnt[lev[i]]++;
        perm[n] = i;
        inv[i] = n;
    }
    // the triangle's dependencies in that numbering (rq / cq), and their consumers
    std::vector<int> rq(nb + 1, 0);
    for (int i = 0; i < nb; ++i) rq[inv[i] + 1] = upper ? rp[i + 1] - diag[i] - 1 : diag[i] - rp[i];
    for (int n = 0; n < nb; ++n) rq[n + 1] += rq[n];
    std::vector<int> cq(rq[nb]), cp(nb + 1, 0);
    for (int i = 0; i < nb; ++i) {
        const int pb = upper ? diag[i] + 1 : rp[i], pe = upper ? rp[i + 1] : diag[i];
        int o = rq[inv[i]];
        for (int p = pb; p < pe; ++p) {
            const int d = inv[ci[p]];
            cq[o++] = d;
            cp[d + 1]++;
        }
    }
    for (int n = 0; n < nb; ++n) cp[n + 1] += cp[n];
    std::vector<int> cons(cp[nb]), cur(cp.begin(), cp.end() - 1), need(nb), unit(nb, -1);
    for (int n = 0; n < nb; ++n) {
        need[n] = rq[n + 1] - rq[n];
        for (int p = rq[n]; p < rq[n + 1]; ++p) cons[cur[cq[p]]++] = n;
    }
    // ready pool: bit n of bits[n / 64], and bit w % 64 of sum[w / 64] when bits[w] != 0
    const int nw = (nb + 63) / 64, ns = (nw + 63) / 64;
    std::vector<uint64_t> bits(nw, 0), sum(ns, 0);
    int low = ns;  // no summary word below it is non-zero
    auto pool_add = [&](int n) {
        bits[n >> 6] |= 1ull << (n & 63);
        sum[n >> 12] |= 1ull << ((n >> 6) & 63);
        low = std::min(low, n >> 12);
    };
    auto pool_pop = [&]() {
        while (!sum[low]) ++low;
        const int w = low * 64 + __builtin_ctzll(sum[low]);
        const int n = w * 64 + __builtin_ctzll(bits[w]);
        bits[w] &= bits[w] - 1;
        if (!bits[w]) sum[low] &= sum[low] - 1;
        return n;
    };
    std::vector<uint64_t> local;  // min-heap of (INT_MAX - inside) << 32 | n
    local.reserve(1024);
    for (int n = 0; n < nb; ++n)
        if (need[n] == 0) pool_add(n);
    order.clear();
    order.reserve(nb);
    for (int u = 0; (int)order.size() < nb; ++u) {
        for (int size = 0; size < usz && (int)order.size() < nb; ++size) {
            int n;
            if (!local.empty()) {
                std::pop_heap(local.begin(), local.end(), std::greater<uint64_t>());
                n = (int)(local.back() & 0xffffffffu);
                local.pop_back();
            } else {
                n = pool_pop();
            }
            unit[n] = u;
            order.push_back(n);
            for (int t = cp[n]; t < cp[n + 1]; ++t) {
                const int c = cons[t];
                if (--need[c] != 0) continue;
                int inside = 0;
                for (int p = rq[c]; p < rq[c + 1]; ++p) inside += unit[cq[p]] == u;
                local.push_back(((uint64_t)(INT_MAX - inside) << 32) | (uint64_t)c);
                std::push_heap(local.begin(), local.end(), std::greater<uint64_t>());
            }
        }
        for (uint64_t e : local) pool_add((int)(e & 0xffffffffu));
        local.clear();
    }
    for (int &n : order) n = perm[n];
}

// f(lo, hi) over [0, n) split into contiguous ranges, one host thread each.
template <class F>
static void parallel_ranges(int n, F f, int grain = 4096);

// Level-synchronous units (bsr4_sptrsv_level_kernel): within each unit of kLvUnit
// consecutive positions of a unit order, the local level of every row (1 + the
// deepest local level of its dependencies in the unit, 0 without), and the unit's
// rows re-sorted by (local level, position).  loc[q] = the local level at position
// q.  Returns false when the kernel cannot take the order: a row with more than
// kLvMaxDeps dependencies, or a wave that would meet its rows out of level order
// (its quads hold slots 16 wv + 128 k + 0..15; the lowest level of its rows k must
// lie above the highest of its rows k - 1).
constexpr int kLvMaxDeps = 3;
static bool level_units(int nb, const std::vector<int> &rp, const std::vector<int> &ci,
                        const std::vector<int> &diag, bool upper, std::vector<int> &order, std::vector<int> &loc) {
    const int nu = (nb + kLvUnit - 1) / kLvUnit;
    std::vector<int> at(nb);
    parallel_ranges(nb, [&](int lo, int hi) {
        for (int q = lo; q < hi; ++q) at[order[q]] = q;
    });
    loc.assign(nb, 0);
    std::atomic<bool> ok{true};
    parallel_ranges(nu, [&](int w0, int w1) {
        std::vector<int> lv(kLvUnit), cnt(kLvUnit + 1), sorted(kLvUnit);
        for (int w = w0; w < w1 && ok.load(std::memory_order_relaxed); ++w) {
            const int base = w * kLvUnit, n = std::min(kLvUnit, nb - base);
            int nl = 0;
            for (int s = 0; s < n; ++s) {  // the unit order is topological
                const int i = order[base + s];
                const int pb = upper ? diag[i] + 1 : rp[i], pe = upper ? rp[i + 1] : diag[i];
                if (pe - pb > kLvMaxDeps) ok.store(false, std::memory_order_relaxed);
                int l = 0;
                for (int p = pb; p < pe; ++p) {
                    const int qc = at[ci[p]];
                    if (qc >= base && qc < base + n) l = std::max(l, lv[qc - base] + 1);
                }
                lv[s] = l;
                nl = std::max(nl, l + 1);
            }
            std::fill(cnt.begin(), cnt.begin() + nl + 1, 0);
            for (int s = 0; s < n; ++s) cnt[lv[s] + 1]++;
            for (int l = 0; l < nl; ++l) cnt[l + 1] += cnt[l];
            for (int s = 0; s < n; ++s) sorted[cnt[lv[s]]++] = s;
            // apply the permutation (the rows of this unit only)
            std::vector<int> rows(n);
            for (int s = 0; s < n; ++s) rows[s] = order[base + sorted[s]];
            for (int s = 0; s < n; ++s) {
                loc[base + s] = lv[sorted[s]];
                order[base + s] = rows[s];
            }
            for (int wv = 0; wv < kLvWaves; ++wv)
                for (int k = 1; k < kLvRows; ++k) {
                    const int first = kLvQuads * k + 16 * wv;
                    if (first >= n) continue;
                    const int prev_last = std::min(kLvQuads * (k - 1) + 16 * wv + 15, n - 1);
                    if (loc[base + first] <= loc[base + prev_last]) ok.store(false, std::memory_order_relaxed);
                }
        }
    }, 8);
    return ok.load();
}

// f(lo, hi) over [0, n) split into contiguous ranges, one host thread each, with at
// least `grain` items per thread (rows: 4096; units, each hundreds of rows: fewer).
template <class F>
static void parallel_ranges(int n, F f, int grain) {
    const int hw = (int)std::thread::hardware_concurrency();
    const int T = std::max(1, std::min({16, hw > 0 ? hw : 1, (n + grain - 1) / grain}));
    if (T == 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 1; t < T; ++t) th.emplace_back(f, (int)((long long)n * t / T), (int)((long long)n * (t + 1) / T));
    f(0, (int)((long long)n / T));
    for (auto &x : th) x.join();
}

}  // namespace isai

extern "C" {

// auto_level: the level kernel was picked by chunk = 0, and is kept only when the
// units' levels are wide enough (kLvMinWidth rows on average): on a 2-D 5-point
// grid a 512-row unit spans about 45 levels of about 11 rows, and the level kernel
// lost to the unit kernel (256^2: 803 against 529 us; 1024^2: 4.49 against
// 2.09 ms), while on 3-D 7-point grids (22 levels of 23 rows) it won at every
// size (32^3: 130 against 174 us; 96^3: 507 against 827 us).
constexpr int kLvMinWidth = 16;
static int analysis_impl(int nb, int nnzb, int bs, const int *d_row, const int *d_col, const double *d_LU,
                         int chunk, bool auto_level, isai_sptrsv_plan_t *plan, isai_stream_t stream) {
    ISAI_REQUIRE(plan, "isai_bsr_sptrsv_analysis: null output");
    *plan = nullptr;
    ISAI_REQUIRE(nb > 0 && nnzb >= nb && d_row && d_col && d_LU, "isai_bsr_sptrsv_analysis: bad argument");
    if (bs != 4) {
        set_error("isai_bsr_sptrsv: the BSR-native solve is built for bs = 4 (got %d); use the CSR form", bs);
        return ISAI_ERR_UNSUPPORTED;
    }
    // chunk < 0: the row schedule; -1 in unit order, -2 in level order, -3 in
    // unit order with the row kernel (the fallback of -1 when a unit has more than
    // kUnitImpMax imports), -4 level-synchronous units (falls back to -1 when the
    // order does not suit bsr4_sptrsv_level_kernel)
    ISAI_REQUIRE(chunk <= kPipeRows && chunk >= -4, "isai_bsr_sptrsv_analysis: chunk %d outside [-4, %d]", chunk,
                 kPipeRows);
    bool rows = chunk < 0;
    const bool units = chunk != -2;
    bool unit_kernel = chunk == -1 || chunk == 0;
    bool level_kernel = chunk == -4;
    int usz = level_kernel ? kLvUnit : kRowsPerWG;  // rows per unit
    const bool automatic = chunk == 0;
    if (rows || automatic) chunk = kPipeRows;
    hipStream_t st = (hipStream_t)stream;
    std::vector<int> rp(nb + 1), ci(nnzb), diag(nb, -1);
    ISAI_HIP(hipMemcpyAsync(rp.data(), d_row, sizeof(int) * (nb + 1), hipMemcpyDeviceToHost, st));
    ISAI_HIP(hipMemcpyAsync(ci.data(), d_col, sizeof(int) * nnzb, hipMemcpyDeviceToHost, st));
    ISAI_HIP(hipStreamSynchronize(st));
    // diagonal positions and pattern checks, over row ranges on host threads; the
    // first offending row is reported
    {
        std::atomic<int> bad_order{INT_MAX}, bad_diag{INT_MAX};
        parallel_ranges(nb, [&](int lo, int hi) {
            for (int i = lo; i < hi; ++i) {
                for (int p = rp[i]; p < rp[i + 1]; ++p) {
                    if (p > rp[i] && ci[p] <= ci[p - 1]) {
                        for (int b = bad_order.load(); i < b && !bad_order.compare_exchange_weak(b, i);) {
                        }
                    }
                    if (ci[p] == i) diag[i] = p;
                }
                if (diag[i] < 0)
                    for (int b = bad_diag.load(); i < b && !bad_diag.compare_exchange_weak(b, i);) {
                    }
            }
        });
        if (bad_order.load() != INT_MAX && bad_order.load() <= bad_diag.load()) {
            set_error("isai_bsr_sptrsv_analysis: block columns of row %d are not strictly ascending", bad_order.load());
            return ISAI_ERR_INVALID_ARG;
        }
        if (bad_diag.load() != INT_MAX) {
            set_error("isai_bsr_sptrsv_analysis: block row %d has no diagonal block", bad_diag.load());
            return ISAI_ERR_ZERO_PIVOT;
        }
    }
    std::vector<int4> chunks[2];
    std::vector<int> rptr[2], ritems[2], lptr[2], litems[2], order[2];
    if (rows || automatic) {
        int nlev = 1;
        std::vector<int> lev[2];
        int nl[2];
        {
            std::thread tu([&] { nl[1] = row_levels(nb, rp, ci, diag, true, lev[1]); });
            nl[0] = row_levels(nb, rp, ci, diag, false, lev[0]);
            tu.join();
        }
        nlev = std::max({nlev, nl[0], nl[1]});
        if (automatic) rows = nb / nlev >= kRowMinWidth;
        if (automatic && rows) {  // the level kernel, kept if its units' levels are wide enough
            level_kernel = auto_level = true;
            unit_kernel = false;
            usz = kLvUnit;
        }
        if (rows && units) {
            // A structurally symmetric pattern makes U's dependency graph L's
            // reversed, so the reversed L order is a U order; when nb is a multiple
            // of the unit size its units are L's units reversed, with the same
            // dependencies inside them.  Otherwise U's order is grown on a second
            // host thread.
            long long low = 0, up = 0;
            for (int i = 0; i < nb; ++i) low += diag[i] - rp[i], up += rp[i + 1] - diag[i] - 1;
            std::atomic<bool> sym{low == up && nb % usz == 0};  // then L = U transposed iff L in U^T
            parallel_ranges(nb, [&](int lo, int hi) {
                for (int i = lo; i < hi && sym.load(std::memory_order_relaxed); ++i)
                    for (int p = rp[i]; p < diag[i]; ++p)
                        if (!std::binary_search(ci.begin() + diag[ci[p]] + 1, ci.begin() + rp[ci[p] + 1], i)) {
                            sym.store(false, std::memory_order_relaxed);
                            break;
                        }
            });
            if (sym.load()) {
                unit_order(nb, rp, ci, diag, false, lev[0], nl[0], usz, order[0]);
                order[1].assign(order[0].rbegin(), order[0].rend());
            } else {
                std::thread tu([&] { unit_order(nb, rp, ci, diag, true, lev[1], nl[1], usz, order[1]); });
                unit_order(nb, rp, ci, diag, false, lev[0], nl[0], usz, order[0]);
                tu.join();
            }
        } else if (rows) {
            for (int u = 0; u < 2; ++u) level_order(nb, lev[u], nl[u], u == 1, order[u]);
        }
    }
    // Level-synchronous units: rows of each unit re-sorted by local level (loc);
    // an order the level kernel cannot take falls back to the unit kernel.
    std::vector<int> loc[2];
    if (level_kernel) {
        bool ok[2];
        std::thread tu([&] { ok[1] = level_units(nb, rp, ci, diag, true, order[1], loc[1]); });
        ok[0] = level_units(nb, rp, ci, diag, false, order[0], loc[0]);
        tu.join();
        if (!ok[0] || !ok[1]) return analysis_impl(nb, nnzb, bs, d_row, d_col, d_LU, -1, false, plan, stream);
        if (auto_level) {
            // rows per local level, averaged over the units of each triangle
            for (int u = 0; u < 2; ++u) {
                long long levels = 0;
                for (int w = 0; w * kLvUnit < nb; ++w) levels += loc[u][std::min(nb, (w + 1) * kLvUnit) - 1] + 1;
                if (nb < (long long)kLvMinWidth * levels)
                    return analysis_impl(nb, nnzb, bs, d_row, d_col, d_LU, -1, false, plan, stream);
            }
        }
    }
    // Row schedules: the position of every row, and the per-plan column array cx.
    // A dependency on a row of the same workgroup (usz consecutive positions: 64,
    // or kLvUnit for the level kernel) is ~slot, read from LDS.  With the unit and
    // level kernels a dependency on another unit is ~(usz + k), k its import slot:
    // the unit's imports are the distinct rows of other units its rows depend on,
    // in order of first use.  One pass per triangle over the units, split over
    // host threads (units are independent).
    std::vector<int> imp_ptr[2], imp_col[2], cx;
    int max_imp = 0;
    if (rows) {
        cx = ci;
        const int nu = (nb + usz - 1) / usz;
        const bool want_imports = units && (unit_kernel || level_kernel);
        for (int u = 0; u < 2; ++u) {
            std::vector<int> at(nb);
            parallel_ranges(nb, [&](int lo, int hi) {
                for (int q = lo; q < hi; ++q) at[order[u][q]] = q;
            });
            // a unit's imports fit in the dependency count of its rows
            std::vector<int> off(nu + 1, 0), cnt(nu, 0), buf;
            if (want_imports) {
                for (int w = 0; w < nu; ++w) {
                    int d = 0;
                    for (int q = w * usz; q < std::min(nb, (w + 1) * usz); ++q) {
                        const int i = order[u][q];
                        d += u == 0 ? diag[i] - rp[i] : rp[i + 1] - diag[i] - 1;
                    }
                    off[w + 1] = off[w] + d;
                }
                buf.resize(std::max(off[nu], 1));
            }
            // a unit's imports are numbered through an open-addressing table of at
            // least twice its dependency count, cleared after the unit at the
            // entries it filled
            int max_d = 0;
            for (int w = 0; want_imports && w < nu; ++w) max_d = std::max(max_d, off[w + 1] - off[w]);
            int hcap = 64;
            while (hcap < 2 * max_d) hcap <<= 1;
            parallel_ranges(nu, [&](int w0, int w1) {
                std::vector<int> hkey, hval, used;
                if (want_imports) {
                    hkey.assign(hcap, -1);
                    hval.resize(hcap);
                }
                for (int w = w0; w < w1; ++w) {
                    int k = 0;
                    unsigned mask = 63;
                    while ((int)mask + 1 < 2 * (off[w + 1] - off[w])) mask = mask * 2 + 1;
                    for (int q = w * usz; q < std::min(nb, (w + 1) * usz); ++q) {
                        const int i = order[u][q];
                        const int pb = u == 0 ? rp[i] : diag[i] + 1, pe = u == 0 ? diag[i] : rp[i + 1];
                        for (int p = pb; p < pe; ++p) {
                            const int c = ci[p];
                            if (at[c] / usz == w) {
                                cx[p] = ~(at[c] % usz);
                            } else if (want_imports) {
                                unsigned h = ((unsigned)c * 2654435761u >> 7) & mask;
                                while (hkey[h] >= 0 && hkey[h] != c) h = (h + 1) & mask;
                                if (hkey[h] < 0) {
                                    hkey[h] = c;
                                    hval[h] = k;
                                    used.push_back((int)h);
                                    buf[off[w] + k++] = c;
                                }
                                cx[p] = ~(usz + hval[h]);
                            }
                        }
                    }
                    for (int h : used) hkey[h] = -1;
                    used.clear();
                    cnt[w] = k;
                }
            }, std::max(1, 4096 / usz));
            if (want_imports) {
                imp_ptr[u].assign(nu + 1, 0);
                for (int w = 0; w < nu; ++w) {
                    imp_ptr[u][w + 1] = imp_ptr[u][w] + cnt[w];
                    max_imp = std::max(max_imp, cnt[w]);
                }
                imp_col[u].resize(imp_ptr[u][nu]);
                parallel_ranges(nu, [&](int w0, int w1) {
                    for (int w = w0; w < w1; ++w)
                        std::copy(buf.begin() + off[w], buf.begin() + off[w] + cnt[w], imp_col[u].begin() + imp_ptr[u][w]);
                }, std::max(1, 4096 / usz));
            }
        }
        if (level_kernel && max_imp > kUnitImpMax)
            return analysis_impl(nb, nnzb, bs, d_row, d_col, d_LU, -1, false, plan, stream);
        if (want_imports && max_imp > kUnitImpMax) {
            // too many imports for the unit kernel's LDS: the row kernel reads
            // dependencies on other units from the solution vector (global columns)
            unit_kernel = false;
            parallel_ranges(nnzb, [&](int lo, int hi) {
                for (int p = lo; p < hi; ++p)
                    if (cx[p] < 0 && ~cx[p] >= kRowsPerWG) cx[p] = ci[p];
            });
        }
    }
    unit_kernel = unit_kernel && rows && units;
    const int row_lds = level_kernel  ? (2 * kLvUnit + max_imp + 1) * 32
                        : unit_kernel ? std::max(kUnitLds, (kRowsPerWG + max_imp) * 32)
                        : units       ? kUnitRowLds
                                      : kRowLds;
    // the smallest batch that holds the longest row of either triangle
    int max_deps = 0;
    for (int i = 0; i < nb; ++i) max_deps = std::max({max_deps, diag[i] - rp[i], rp[i + 1] - diag[i] - 1});
    const int unit_batch = max_deps <= 2 ? 2 : max_deps <= 4 ? 4 : 8;
    const int lv_batch = max_deps <= 2 ? 2 : 3;  // level kernel: at most kLvMaxDeps, checked above
    if (level_kernel) {
        ISAI_HIP(lv_batch == 2 ? level_attr<2>(row_lds) : level_attr<3>(row_lds));
    } else if (unit_kernel) {
        ISAI_HIP(unit_batch == 2 ? unit_attr<2>(row_lds) : unit_batch == 4 ? unit_attr<4>(row_lds) : unit_attr<8>(row_lds));
    } else if (rows) {
        ISAI_HIP(hipFuncSetAttribute((const void *)bsr4_sptrsv_row_kernel<false>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, row_lds));
        ISAI_HIP(hipFuncSetAttribute((const void *)bsr4_sptrsv_row_kernel<true>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, row_lds));
    }
    for (int u = 0; u < 2 && !rows; ++u) {
        int s = build_tables(nb, rp, ci, diag, u == 1, chunk, chunks[u], rptr[u], ritems[u], lptr[u], litems[u]);
        if (s != ISAI_OK) return s;
    }
    auto *P = new isai_sptrsv_plan;
    P->rows = rows;
    P->units = rows && units;
    P->unit_kernel = unit_kernel;
    P->unit_batch = unit_batch;
    P->level_kernel = level_kernel;
    P->lv_batch = lv_batch;
    P->row_lds = row_lds;
    P->nb = nb;
    P->nnzb = nnzb;
    P->bs = bs;
    P->chunk = chunk;
    P->d_row = d_row;
    P->d_col = d_col;
    bool ok = hipMalloc(&P->d_diag, sizeof(int) * nb) == hipSuccess &&
              hipMalloc(&P->d_dinvL, sizeof(double) * 16 * nb) == hipSuccess &&
              hipMalloc(&P->d_dinvU, sizeof(double) * 16 * nb) == hipSuccess &&
              hipMalloc(&P->d_y, sizeof(double) * 4 * nb) == hipSuccess &&
              hipMalloc(&P->d_S, sizeof(double) * 16 * (size_t)nnzb) == hipSuccess &&
              hipMalloc(&P->d_cnt, sizeof(Counters) + sizeof(int)) == hipSuccess;
    for (int u = 0; u < 2; ++u) {
        P->nchunks[u] = (int)chunks[u].size() / 2;
        for (int c = 0; c < P->nchunks[u]; ++c) {
            const int4 a = chunks[u][2 * c], b = chunks[u][2 * c + 1];
            P->maxr[u] = std::max(P->maxr[u], a.y - a.x);
            P->maxrem[u] = std::max(P->maxrem[u], b.x - a.z);
            P->maxloc[u] = std::max(P->maxloc[u], b.y - a.w);
        }
    }
    for (int u = 0; u < 2 && ok && rows; ++u) ok = hipMalloc(&P->d_info[u], sizeof(int4) * nb) == hipSuccess;
    if (ok && rows) ok = hipMalloc(&P->d_cx, sizeof(int) * (size_t)nnzb) == hipSuccess;
    for (int u = 0; u < 2 && ok && (unit_kernel || level_kernel); ++u)
        ok = hipMalloc(&P->d_imp_ptr[u], sizeof(int) * imp_ptr[u].size()) == hipSuccess &&
             hipMalloc(&P->d_imp_col[u], sizeof(int) * std::max<size_t>(imp_col[u].size(), 1)) == hipSuccess;
    for (int u = 0; u < 2 && ok && !rows; ++u) {
        ok = hipMalloc(&P->d_chunks[u], sizeof(int4) * chunks[u].size()) == hipSuccess &&
             hipMalloc(&P->d_rem_ptr[u], sizeof(int) * (nb + 1)) == hipSuccess &&
             hipMalloc(&P->d_loc_ptr[u], sizeof(int) * (nb + 1)) == hipSuccess &&
             hipMalloc(&P->d_rem_items[u], sizeof(int) * std::max<size_t>(ritems[u].size(), 1)) == hipSuccess &&
             hipMalloc(&P->d_loc_items[u], sizeof(int) * std::max<size_t>(litems[u].size(), 1)) == hipSuccess;
    }
    if (!ok) {
        (void)hipGetLastError();
        isai_bsr_sptrsv_destroy(P);
        set_error("isai_bsr_sptrsv_analysis: hipMalloc failed");
        return ISAI_ERR_ALLOC;
    }
    // uploads; the plan is released if one fails
    auto upload = [&]() -> int {
        ISAI_HIP(hipMemcpy(P->d_diag, diag.data(), sizeof(int) * nb, hipMemcpyHostToDevice));
        if (rows) {
            std::vector<int4> info(nb);
            for (int u = 0; u < 2; ++u) {
                parallel_ranges(nb, [&](int lo, int hi) {
                    for (int q = lo; q < hi; ++q) {
                        const int i = order[u][q];
                        const int w4 = level_kernel ? loc[u][q] : 1;  // local level (level kernel)
                        info[q] = u == 0 ? make_int4(i, rp[i], diag[i], w4) : make_int4(i, diag[i] + 1, rp[i + 1], w4);
                    }
                });
                ISAI_HIP(hipMemcpy(P->d_info[u], info.data(), sizeof(int4) * nb, hipMemcpyHostToDevice));
                if (unit_kernel || level_kernel) {
                    ISAI_HIP(hipMemcpy(P->d_imp_ptr[u], imp_ptr[u].data(), sizeof(int) * imp_ptr[u].size(),
                                       hipMemcpyHostToDevice));
                    if (!imp_col[u].empty())
                        ISAI_HIP(hipMemcpy(P->d_imp_col[u], imp_col[u].data(), sizeof(int) * imp_col[u].size(),
                                           hipMemcpyHostToDevice));
                }
            }
            ISAI_HIP(hipMemcpy(P->d_cx, cx.data(), sizeof(int) * (size_t)nnzb, hipMemcpyHostToDevice));
        }
        for (int u = 0; u < 2 && !rows; ++u) {
            ISAI_HIP(hipMemcpy(P->d_chunks[u], chunks[u].data(), sizeof(int4) * chunks[u].size(), hipMemcpyHostToDevice));
            ISAI_HIP(hipMemcpy(P->d_rem_ptr[u], rptr[u].data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice));
            ISAI_HIP(hipMemcpy(P->d_loc_ptr[u], lptr[u].data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice));
            if (!ritems[u].empty())
                ISAI_HIP(hipMemcpy(P->d_rem_items[u], ritems[u].data(), sizeof(int) * ritems[u].size(),
                                   hipMemcpyHostToDevice));
            if (!litems[u].empty())
                ISAI_HIP(hipMemcpy(P->d_loc_items[u], litems[u].data(), sizeof(int) * litems[u].size(),
                                   hipMemcpyHostToDevice));
        }
        return ISAI_OK;
    };
    if (const int us = upload(); us != ISAI_OK) {
        isai_bsr_sptrsv_destroy(P);
        return us;
    }
    Counters h{};
    // steps without progress before a chunk gives up on its missing inputs: a
    // step is one agent-scope poll round (about a microsecond), so this bound is
    // tens of milliseconds, far above any legitimate wait inside one solve.
    h.spin_limit = 1 << 16;
    const hipError_t e = hipMemcpy(P->d_cnt, &h, sizeof(h), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        isai_bsr_sptrsv_destroy(P);
        return hip_fail(e, "hipMemcpy", __FILE__, __LINE__);
    }
    const int us = isai_bsr_sptrsv_update_values(P, d_LU, stream);
    if (us != ISAI_OK) {
        isai_bsr_sptrsv_destroy(P);
        return us;
    }
    *plan = P;
    return ISAI_OK;
}

int isai_bsr_sptrsv_analysis(int nb, int nnzb, int bs, const int *d_row, const int *d_col, const double *d_LU,
                             int chunk, isai_sptrsv_plan_t *plan, isai_stream_t stream) {
    return analysis_impl(nb, nnzb, bs, d_row, d_col, d_LU, chunk, false, plan, stream);
}

int isai_bsr_sptrsv_update_values(isai_sptrsv_plan_t P, const double *d_LU, isai_stream_t stream) {
    ISAI_REQUIRE(P && d_LU, "isai_bsr_sptrsv_update_values: null argument");
    int s = isai_build_dinv(P->nb, 4, P->d_row, P->d_col, d_LU, 1, P->d_dinvL, stream);
    if (s == ISAI_OK) s = isai_build_dinv(P->nb, 4, P->d_row, P->d_col, d_LU, 0, P->d_dinvU, stream);
    if (s != ISAI_OK) return s;
    scale_offdiag_kernel<<<grid_for(P->nb, kScaleBlock / 64), kScaleBlock, 0, (hipStream_t)stream>>>(
        P->nb, P->d_row, P->d_diag, d_LU, P->d_dinvL, P->d_dinvU, P->d_S);
    ISAI_LAUNCH_CHECK();
    return ISAI_OK;
}

int isai_bsr_sptrsv_solve(isai_sptrsv_plan_t P, const double *d_LU, const double *d_b, double *d_x,
                          isai_stream_t stream) {
    ISAI_REQUIRE(P && d_LU && d_b && d_x, "isai_bsr_sptrsv_solve: null argument");
    return isai::sptrsv_solve_internal(P, d_LU, d_b, d_x, (hipStream_t)stream, nullptr);
}

int isai_bsr_sptrsv_set_spin_limit(isai_sptrsv_plan_t P, int limit) {
    ISAI_REQUIRE(P && limit > 0, "isai_bsr_sptrsv_set_spin_limit: bad argument");
    ISAI_HIP(hipMemcpy(&P->d_cnt->spin_limit, &limit, sizeof(int), hipMemcpyHostToDevice));
    return ISAI_OK;
}

int isai_sptrsv_status(isai_sptrsv_plan_t P, isai_stream_t stream) {
    ISAI_REQUIRE(P, "isai_sptrsv_status: null plan");
    hipStream_t st = (hipStream_t)stream;
    int err = 0;
    ISAI_HIP(hipMemcpyAsync(&err, &P->d_cnt->err, sizeof(int), hipMemcpyDeviceToHost, st));
    ISAI_HIP(hipStreamSynchronize(st));
    if (err) {
        const int zero = 0;
        ISAI_HIP(hipMemcpy(&P->d_cnt->err, &zero, sizeof(int), hipMemcpyHostToDevice));
        set_error("sptrsv: a dependency wait exceeded its spin limit (malformed pattern or missing diagonal)");
        return ISAI_ERR_TIMEOUT;
    }
    return ISAI_OK;
}

int isai_bsr_sptrsv_schedule(isai_sptrsv_plan_t P, int *schedule) {
    ISAI_REQUIRE(P && schedule, "isai_bsr_sptrsv_schedule: null argument");
    *schedule = P->level_kernel ? -4 : P->unit_kernel ? -1 : P->rows ? (P->units ? -3 : -2) : P->chunk;
    return ISAI_OK;
}

int isai_bsr_sptrsv_destroy(isai_sptrsv_plan_t P) {
    if (!P) return ISAI_OK;
    (void)hipFree(P->d_diag);
    (void)hipFree(P->d_dinvL);
    (void)hipFree(P->d_dinvU);
    (void)hipFree(P->d_y);
    (void)hipFree(P->d_S);
    (void)hipFree(P->d_cnt);
    (void)hipFree(P->d_cx);
    for (int u = 0; u < 2; ++u) {
        (void)hipFree(P->d_imp_ptr[u]);
        (void)hipFree(P->d_imp_col[u]);
    }
    for (int u = 0; u < 2; ++u) {
        (void)hipFree(P->d_info[u]);
        (void)hipFree(P->d_chunks[u]);
        (void)hipFree(P->d_rem_ptr[u]);
        (void)hipFree(P->d_rem_items[u]);
        (void)hipFree(P->d_loc_ptr[u]);
        (void)hipFree(P->d_loc_items[u]);
    }
    delete P;
    return ISAI_OK;
}

int isai_csr_sptrsv_lu_device(int m, int nnz, const int *d_rp, const int *d_ci, const double *d_v,
                              const double *d_b, double *d_x, void *d_work, int chunk, int spin_limit,
                              isai_stream_t stream) {
    ISAI_REQUIRE(m > 0 && nnz >= 0 && d_rp && d_ci && d_v && d_b && d_x && d_work,
                 "isai_csr_sptrsv_lu_device: bad argument");
    if (chunk <= 0) chunk = 256;
    ISAI_REQUIRE(chunk <= kCsrMaxChunk, "isai_csr_sptrsv_lu_device: chunk above %d", kCsrMaxChunk);
    hipStream_t st = (hipStream_t)stream;
    const int nchunks = (m + chunk - 1) / chunk;
    // d_work: m doubles for the intermediate vector, then a Counters record.
    double *y = (double *)d_work;
    Counters *cnt = (Counters *)((double *)d_work + ((size_t)m + 7) / 8 * 8);
    Counters h{};
    h.spin_limit = spin_limit > 0 ? spin_limit : (1 << 20);
    ISAI_HIP(hipMemcpyAsync(cnt, &h, sizeof(h), hipMemcpyHostToDevice, st));
    fill_sentinel_kernel<<<fill_blocks(m), 256, 0, st>>>(y, (size_t)m, nullptr);
    csr_sptrsv_kernel<false><<<nchunks, 64, 0, st>>>(m, chunk, nchunks, d_rp, d_ci, d_v, d_b, y, d_x, cnt);
    csr_sptrsv_kernel<true><<<nchunks, 64, 0, st>>>(m, chunk, nchunks, d_rp, d_ci, d_v, y, d_x, nullptr, cnt);
    ISAI_LAUNCH_CHECK();
    int err = 0;
    ISAI_HIP(hipMemcpyAsync(&err, &cnt->err, sizeof(int), hipMemcpyDeviceToHost, st));
    ISAI_HIP(hipStreamSynchronize(st));
    if (err) {
        set_error("csr sptrsv: a dependency wait exceeded its spin limit (missing diagonal or malformed pattern)");
        return ISAI_ERR_TIMEOUT;
    }
    return ISAI_OK;
}

int isai_csr_sptrsv_work_size(int m, size_t *bytes) {
    ISAI_REQUIRE(m >= 0 && bytes, "isai_csr_sptrsv_work_size: bad argument");
    *bytes = sizeof(double) * (((size_t)m + 7) / 8 * 8) + sizeof(Counters);
    return ISAI_OK;
}

int isai_csr_sptrsv_lu(int m, const int *rowptr, const int *colind, const double *val, int nnz, double *b) {
    ISAI_REQUIRE(m > 0 && nnz >= 0 && rowptr && colind && val && b, "isai_csr_sptrsv_lu: bad argument");
    int *d_rp = nullptr, *d_ci = nullptr;
    double *d_v = nullptr, *d_b = nullptr, *d_x = nullptr, *d_w = nullptr;
    size_t work = 0;
    int s = isai_csr_sptrsv_work_size(m, &work);
    if (s != ISAI_OK) return s;
    if (hipMalloc(&d_rp, sizeof(int) * (m + 1)) != hipSuccess || hipMalloc(&d_ci, sizeof(int) * (nnz ? nnz : 1)) != hipSuccess ||
        hipMalloc(&d_v, sizeof(double) * (nnz ? nnz : 1)) != hipSuccess || hipMalloc(&d_b, sizeof(double) * m) != hipSuccess ||
        hipMalloc(&d_x, sizeof(double) * m) != hipSuccess ||
        hipMalloc(&d_w, work) != hipSuccess) {
        set_error("isai_csr_sptrsv_lu: hipMalloc failed");
        s = ISAI_ERR_ALLOC;
    }
    if (s == ISAI_OK) {
        hipError_t e = hipMemcpy(d_rp, rowptr, sizeof(int) * (m + 1), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d_ci, colind, sizeof(int) * nnz, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d_v, val, sizeof(double) * nnz, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d_b, b, sizeof(double) * m, hipMemcpyHostToDevice);
        if (e != hipSuccess) s = hip_fail(e, "hipMemcpy", __FILE__, __LINE__);
    }
    if (s == ISAI_OK) s = isai_csr_sptrsv_lu_device(m, nnz, d_rp, d_ci, d_v, d_b, d_x, d_w, 0, 0, nullptr);
    if (s == ISAI_OK) {
        hipError_t e = hipMemcpy(b, d_x, sizeof(double) * m, hipMemcpyDeviceToHost);
        if (e != hipSuccess) s = hip_fail(e, "hipMemcpy", __FILE__, __LINE__);
    }
    (void)hipFree(d_rp);
    (void)hipFree(d_ci);
    (void)hipFree(d_v);
    (void)hipFree(d_b);
    (void)hipFree(d_x);
    (void)hipFree(d_w);
    return s;
}

}  // extern "C"
