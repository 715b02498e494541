// Optimal-transport refinement: the reference's random pair-swap descent
// (thread::morph / morph_asynch, thread.cpp:989-1068) on gfx950.
//
// Acceptance.  The reference swaps p[x1][y] and p[x2][y] when c1 >= c2
// (ties accepted), with c1/c2 the four squared distances to rows y-1 and
// y+1 before/after (thread.cpp:1014-1022; exact integers summed in double).
// Writing C for a slot's point in row y and S = P + N for the sum of its
// neighbours in rows y-1 and y+1 (fixed while row y is refined), every
// squared norm cancels and
//        c1 - c2 = -2 (C1 - C2) . (S1 - S2),
// so a swap is taken iff (C1 - C2).(S1 - S2) <= 0.  With 16.8 fixed point
// (|C| < 2^24, |S| < 2^25) the dot product fits int64 exactly, so every
// decision equals the reference's.
//
// Schedule.  A round is a random (near-)perfect matching of the slots of
// row y: the fold {i, c - i mod w} for a random c (round_params in
// common.h).  Pairs of a round are disjoint and only read the neighbour
// rows, so all of them run in parallel with no conflicts and the round
// equals the same attempts made one after another.
#include "common.h"
#include "../../include/atomorph_hip.h"

#include <algorithm>
#include <vector>

using namespace amh;

namespace {

constexpr int kCostThreads = 256;
constexpr int kRoundThreads = 256;

// ------------------------------------------------------------ chain cost
__global__ __launch_bounds__(kCostThreads) void k_chain_cost(const uint64_t *__restrict__ pts, int h, int64_t wtot,
                                                             const int64_t *__restrict__ off,
                                                             const int32_t *__restrict__ cw, int nchains,
                                                             unsigned long long *cost) {
    const int64_t a = int64_t(blockIdx.x) * kCostThreads + threadIdx.x;
    const bool valid = a < wtot;
    int c = 0;
    uint64_t v = 0;
    if (valid) {
        int lo = 0, hi = nchains - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (off[mid] <= a) lo = mid; else hi = mid - 1;
        }
        c = lo;
        if (cw[c] > 1) {
            uint64_t prev = pts[int64_t(h - 1) * wtot + a];
            for (int j = 0; j < h; ++j) {
                uint64_t cur = pts[int64_t(j) * wtot + a];
                v += pt_dist(prev, cur);
                prev = cur;
            }
        }
    }
    const int c0 = __shfl(c, 0, 64);
    const bool uniform = __all(!valid || c == c0);
    if (uniform) {
        uint64_t s = wave_sum_u64(v);
        if ((threadIdx.x & 63) == 0 && s)
            atomic_add_u128(&cost[2 * c0], &cost[2 * c0 + 1], s);
    } else if (valid && v) {
        atomic_add_u128(&cost[2 * c], &cost[2 * c + 1], v);
    }
}

// The counters of a workgroup, summed over its waves first: one set of
// atomics per workgroup on the three shared counters (device atomics to one
// address serialise, so a set per wave cost as much as the refinement in a
// pass).  Every thread of the workgroup calls it.
template <int NT>
__device__ __forceinline__ void flush_stats(amh_refine_stats *st, uint64_t att, uint64_t acc, int64_t gain) {
    __shared__ uint64_t part[NT / 64][3];
    att = wave_sum_u64(att);
    acc = wave_sum_u64(acc);
    gain = wave_sum_i64(gain);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        part[wave][0] = att;
        part[wave][1] = acc;
        part[wave][2] = uint64_t(gain);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t a = 0, b = 0, g = 0;
        for (int w = 0; w < NT / 64; ++w) {
            a += part[w][0];
            b += part[w][1];
            g += part[w][2];
        }
        if (a) {
            atomicAdd((unsigned long long *)&st->attempts, (unsigned long long)a);
            atomicAdd((unsigned long long *)&st->accepted, (unsigned long long)b);
            atomicAdd((unsigned long long *)&st->gain, (unsigned long long)g);
        }
    }
}

// One pair of one round.  Returns the signed dot product d; the caller swaps
// when d <= 0 (c1 >= c2) and records a gain of -2d.
__device__ __forceinline__ int64_t swap_test(uint64_t ci, uint64_t cj, int32_t sxi, int32_t syi, int32_t sxj,
                                             int32_t syj) {
    return int64_t(pt_fx(ci) - pt_fx(cj)) * int64_t(sxi - sxj) +
           int64_t(pt_fy(ci) - pt_fy(cj)) * int64_t(syi - syj);
}

// -------------------------------------------- plain: one round, one chain
__global__ __launch_bounds__(kRoundThreads) void k_refine_round(uint64_t *__restrict__ pts, int h, int64_t wtot,
                                                                int64_t off, int32_t w, int32_t y, RoundPar rp,
                                                                amh_refine_stats *st) {
    const uint32_t t = blockIdx.x * kRoundThreads + threadIdx.x;
    uint64_t att = 0, acc = 0;
    int64_t gain = 0;
    if (t < rp.npairs) {
        uint32_t i, j;
        round_pair(rp, uint32_t(w), t, i, j);
        {
            const int yp = y > 0 ? y - 1 : h - 1, yn = (y + 1) % h;
            uint64_t *row = pts + int64_t(y) * wtot + off;
            const uint64_t *rprev = pts + int64_t(yp) * wtot + off;
            const uint64_t *rnext = pts + int64_t(yn) * wtot + off;
            const uint64_t ci = row[i], cj = row[j];
            const uint64_t pi = rprev[i], pj = rprev[j], ni = rnext[i], nj = rnext[j];
            const int64_t d = swap_test(ci, cj, pt_fx(pi) + pt_fx(ni), pt_fy(pi) + pt_fy(ni), pt_fx(pj) + pt_fx(nj),
                                        pt_fy(pj) + pt_fy(nj));
            att = 1;
            if (d <= 0) {
                row[i] = cj;
                row[j] = ci;
                acc = 1;
                gain = -2 * d;
            }
        }
    }
    flush_stats<kRoundThreads>(st, att, acc, gain);
}

// ---------------------------- plain: a whole visit of one chain, one launch
// All `rounds` rounds of one visit in one workgroup, for chains of up to
// kVisitCap slots: row y and the slots' neighbour sums are staged in LDS,
// the rounds run there (they are sequential: each reads the swaps of the
// one before, hence a barrier between rounds), and the row is written back
// once.  The same pairs in the same order as one k_refine_round launch per
// round.
constexpr int kVisitThreads = 1024;
constexpr int kVisitCap = 8192;

__global__ __launch_bounds__(kVisitThreads) void k_refine_visit(uint64_t *__restrict__ pts, int h, int64_t wtot,
                                                                int64_t off, int32_t w, int32_t y, uint64_t seed,
                                                                uint64_t stream_id, uint64_t rounds,
                                                                amh_refine_stats *st) {
    __shared__ uint64_t C[kVisitCap];
    __shared__ int2 S[kVisitCap];  // neighbour sums (x, y) of each slot
    const int yp = y > 0 ? y - 1 : h - 1, yn = (y + 1) % h;
    uint64_t *row = pts + int64_t(y) * wtot + off;
    const uint64_t *rprev = pts + int64_t(yp) * wtot + off;
    const uint64_t *rnext = pts + int64_t(yn) * wtot + off;
    for (int l = threadIdx.x; l < w; l += kVisitThreads) {
        const uint64_t p = rprev[l], q = rnext[l];
        C[l] = row[l];
        S[l] = make_int2(pt_fx(p) + pt_fx(q), pt_fy(p) + pt_fy(q));
    }
    __syncthreads();
    uint64_t att = 0, acc = 0;
    int64_t gain = 0;
    for (uint64_t r = 0; r < rounds; ++r) {
        const RoundPar rp = round_params(seed, stream_id, r, uint32_t(w));
        for (uint32_t t = threadIdx.x; t < rp.npairs; t += kVisitThreads) {
            uint32_t i, j;
            round_pair(rp, uint32_t(w), t, i, j);
            const uint64_t ci = C[i], cj = C[j];
            const int2 si = S[i], sj = S[j];
            const int64_t d = swap_test(ci, cj, si.x, si.y, sj.x, sj.y);
            ++att;
            if (d <= 0) {
                C[i] = cj;
                C[j] = ci;
                ++acc;
                gain += -2 * d;
            }
        }
        __syncthreads();
    }
    for (int l = threadIdx.x; l < w; l += kVisitThreads) row[l] = C[l];
    flush_stats<kVisitThreads>(st, att, acc, gain);
}

// ------------------------------------- batched: persistent, LDS resident
// queue[0] = next item, queue[1] = workgroups finished, queue[2] = pass
// counter.  The last workgroup to finish resets the first two and advances
// the counter, so the kernel can be replayed (hipGraph) without a memset
// node; with pass == kPassFromQueue every replay runs the next pass.
constexpr uint64_t kPassFromQueue = ~0ull;

template <int CAP, int NT>
__global__ __launch_bounds__(NT) void k_refine_pass(uint64_t *__restrict__ pts, int h, int64_t wtot,
                                                    const int64_t *__restrict__ off, const int32_t *__restrict__ cw,
                                                    const int32_t *__restrict__ items, int nitems, uint64_t pass,
                                                    int rounds, uint64_t seed, uint32_t *queue,
                                                    amh_refine_stats *st) {
    __shared__ uint64_t C[CAP];
    __shared__ int2 S[CAP];  // neighbour sums (x, y), one 8-byte LDS load per slot
    __shared__ int item_s;
    __shared__ uint32_t pass_s;
    const int tid = threadIdx.x;
    if (pass == kPassFromQueue) {
        if (tid == 0) pass_s = __hip_atomic_load(&queue[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        pass = pass_s;
    }
    uint64_t att = 0, acc = 0;
    int64_t gain = 0;
    // With a workgroup per item the items are taken in launch order;
    // otherwise they come from the queue, largest first.
    const bool one_each = nitems <= int(gridDim.x);
    for (int turn = 0;; ++turn) {
        if (one_each) {
            if (tid == 0) item_s = turn == 0 ? int(blockIdx.x) : nitems;
        } else if (tid == 0) {
            item_s = int(atomicAdd(&queue[0], 1u));
        }
        __syncthreads();
        const int it = item_s;
        if (it >= nitems) break;
        const int c = items[2 * it], g = items[2 * it + 1];
        const int32_t w = cw[c];
        const int64_t o = off[c];
        const int y = int(pass_row(seed, uint64_t(c), pass, uint32_t(h)));
        const int yp = y > 0 ? y - 1 : h - 1, yn = (y + 1) % h;
        const int ng = (w + CAP - 1) / CAP;
        const int gs = (w + ng - 1) / ng;
        const int g0 = g * gs;
        const int n = min(gs, w - g0);
        const bool permuted = ng > 1;
        const uint64_t pkey = pass_stream(uint64_t(c), pass, 0xFFFFFFFFull);
        const uint64_t *rprev = pts + int64_t(yp) * wtot + o;
        const uint64_t *rnext = pts + int64_t(yn) * wtot + o;
        uint64_t *row = pts + int64_t(y) * wtot + o;
        for (int l = tid; l < n; l += NT) {
            const uint32_t s = permuted ? feistel_perm(uint32_t(g0 + l), uint32_t(w), pkey) : uint32_t(l);
            const uint64_t p = rprev[s], q = rnext[s];
            C[l] = row[s];
            S[l] = make_int2(pt_fx(p) + pt_fx(q), pt_fy(p) + pt_fy(q));
        }
        __syncthreads();
        if (n >= 2) {
            const uint64_t sid = pass_stream(uint64_t(c), pass, uint64_t(g));
            for (int r = 0; r < rounds; ++r) {
                const RoundPar rp = round_params(seed, sid, uint64_t(r), uint32_t(n));
                // A thread's pairs of the round (at most kPer): the pairs are
                // disjoint, so all their loads are issued before any swap.
                constexpr int kPer = (CAP / 2 + NT - 1) / NT;
                uint32_t pi[kPer], pj[kPer];
                uint64_t ci[kPer], cj[kPer];
                int2 si[kPer], sj[kPer];
#pragma unroll
                for (int k = 0; k < kPer; ++k) {
                    const uint32_t t = uint32_t(tid + k * NT);
                    if (t < rp.npairs) {
                        round_pair(rp, uint32_t(n), t, pi[k], pj[k]);
                        ci[k] = C[pi[k]];
                        cj[k] = C[pj[k]];
                        si[k] = S[pi[k]];
                        sj[k] = S[pj[k]];
                    }
                }
#pragma unroll
                for (int k = 0; k < kPer; ++k) {
                    if (uint32_t(tid + k * NT) >= rp.npairs) continue;
                    const int64_t d = swap_test(ci[k], cj[k], si[k].x, si[k].y, sj[k].x, sj[k].y);
                    ++att;
                    if (d <= 0) {
                        C[pi[k]] = cj[k];
                        C[pj[k]] = ci[k];
                        ++acc;
                        gain += -2 * d;
                    }
                }
                __syncthreads();
            }
        }
        for (int l = tid; l < n; l += NT) {
            const uint32_t s = permuted ? feistel_perm(uint32_t(g0 + l), uint32_t(w), pkey) : uint32_t(l);
            row[s] = C[l];
        }
        __syncthreads();
    }
    flush_stats<NT>(st, att, acc, gain);
    if (tid == 0) {
        __threadfence();
        const uint32_t done = atomicAdd(&queue[1], 1u);
        if (done == gridDim.x - 1) {
            atomicExch(&queue[0], 0u);
            atomicExch(&queue[1], 0u);
            atomicAdd(&queue[2], 1u);
        }
    }
}

// ------------------------------------------- batched, kNN-guided pairs
// The reference's OpenCV build draws the second slot of a swap from a
// nearest-neighbour search instead of at random (thread.cpp:893-986): for
// slot x1 of row y it takes the two points of row y+1 nearest to x1's point
// in row y (slots whose row-(y+1) point is where x1's point would like to
// go), the two nearest to x1's own point in row y+1, and one random slot,
// and picks one of them at random.
//
// Here every slot of a work item proposes one partner per round that way,
// over an index of row y+1 built once per item in LDS: the slots sorted by
// the Morton key of their row-(y+1) pixel.  A point's two nearest are taken
// from the four entries around its place in that order -- an approximate
// search, as the reference's FLANN search with one check is.  Proposals
// overlap, so each claims both of its slots with its priority (atomicMin in
// LDS) and only those holding both claims run their swap test: the winners'
// pairs are disjoint, so they run in parallel and equal the same attempts
// made one after another.  Row y+1 is not permuted while row y is refined,
// so the index holds for the whole pass.
constexpr uint32_t kNoClaim = 0xFFFFFFFFu;

// The two entries of sk[lo, hi) (clipped to [0, n)) whose pixels are
// nearest to (qx, qy), other than slot `self`; ties go to the earlier entry.
__device__ __forceinline__ void knn_nearest2(const uint64_t *sk, int n, int lo, int hi, uint32_t self, uint32_t qx,
                                             uint32_t qy, uint32_t *cand, int &m) {
    uint64_t d1 = ~0ull, d2 = ~0ull;
    int b1 = -1, b2 = -1;
    for (int k = max(lo, 0); k < min(hi, n); ++k) {
        const uint64_t e = sk[k];
        const uint32_t s = uint32_t(e & 0xFFFFu);
        if (s == self) continue;
        const uint32_t key = uint32_t(e >> 16);
        const int64_t dx = int64_t(squeeze16(key)) - qx, dy = int64_t(squeeze16(key >> 1)) - qy;
        const uint64_t dd = uint64_t(dx * dx + dy * dy);
        if (dd < d1) {
            d2 = d1;
            b2 = b1;
            d1 = dd;
            b1 = int(s);
        } else if (dd < d2) {
            d2 = dd;
            b2 = int(s);
        }
    }
    if (b1 >= 0) cand[m++] = uint32_t(b1);
    if (b2 >= 0) cand[m++] = uint32_t(b2);
}

// The partner slot slot l proposes with draw d (knn_draw).
__device__ __forceinline__ uint32_t knn_partner(const uint64_t *sk, const uint16_t *rk, const uint64_t *cur, int n,
                                                uint32_t l, uint64_t d) {
    uint32_t cand[5];
    int m = 0;
    // near the slot's point in row y: where it would like to go
    const uint64_t q = cur[l];
    const uint32_t qx = pt_x(q), qy = pt_y(q);
    const uint64_t kq = uint64_t(morton2(qx, qy)) << 16;
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (sk[mid] < kq) lo = mid + 1;
        else hi = mid;
    }
    knn_nearest2(sk, n, lo - 2, lo + 2, l, qx, qy, cand, m);
    // near the slot's own point in row y+1
    const int kd = rk[l];
    const uint32_t key = uint32_t(sk[kd] >> 16);
    const uint32_t nx = squeeze16(key), ny = squeeze16(key >> 1);
    knn_nearest2(sk, n, kd - 2, kd + 3, l, nx, ny, cand, m);
    // and one random other slot
    const uint32_t u = uint32_t((((d >> 32) & 0xFFFFu) * uint64_t(n - 1)) >> 16);
    cand[m++] = u < l ? u : u + 1u;
    return cand[((d >> 48) * uint64_t(m)) >> 16];
}

template <int CAP, int NT>
__global__ __launch_bounds__(NT) void k_refine_knn(uint64_t *__restrict__ pts, int h, int64_t wtot,
                                                   const int64_t *__restrict__ off, const int32_t *__restrict__ cw,
                                                   const int32_t *__restrict__ items, int nitems, uint64_t pass,
                                                   int rounds, uint64_t seed, uint32_t *queue,
                                                   amh_refine_stats *st) {
    static_assert(CAP <= 4096 && (CAP & (CAP - 1)) == 0 && CAP % NT == 0, "slots in 12 bits, whole rounds");
    constexpr int PER = CAP / NT;
    __shared__ uint64_t C[CAP];
    __shared__ int32_t SX[CAP], SY[CAP];
    __shared__ uint64_t SK[CAP];  // Morton key of the row-(y+1) pixel << 16 | slot, sorted
    __shared__ uint16_t RK[CAP];  // place of each slot in SK
    __shared__ uint32_t CL[CAP];  // claims of the round
    __shared__ int item_s;
    const int tid = threadIdx.x;
    uint64_t att = 0, acc = 0;
    int64_t gain = 0;
    const bool one_each = nitems <= int(gridDim.x);
    for (int turn = 0;; ++turn) {
        if (one_each) {
            if (tid == 0) item_s = turn == 0 ? int(blockIdx.x) : nitems;
        } else if (tid == 0) {
            item_s = int(atomicAdd(&queue[0], 1u));
        }
        __syncthreads();
        const int it = item_s;
        if (it >= nitems) break;
        const int c = items[2 * it], g = items[2 * it + 1];
        const int32_t w = cw[c];
        const int64_t o = off[c];
        const int y = int(pass_row(seed, uint64_t(c), pass, uint32_t(h)));
        const int yp = y > 0 ? y - 1 : h - 1, yn = (y + 1) % h;
        const int ng = (w + CAP - 1) / CAP;
        const int gs = (w + ng - 1) / ng;
        const int g0 = g * gs;
        const int n = min(gs, w - g0);
        const bool permuted = ng > 1;
        const uint64_t pkey = pass_stream(uint64_t(c), pass, 0xFFFFFFFFull);
        const uint64_t *rprev = pts + int64_t(yp) * wtot + o;
        const uint64_t *rnext = pts + int64_t(yn) * wtot + o;
        uint64_t *row = pts + int64_t(y) * wtot + o;
        int P = 2;
        while (P < n) P <<= 1;
        for (int l = tid; l < P; l += NT) {
            if (l < n) {
                const uint32_t s = permuted ? feistel_perm(uint32_t(g0 + l), uint32_t(w), pkey) : uint32_t(l);
                const uint64_t p = rprev[s], q = rnext[s];
                C[l] = row[s];
                SX[l] = pt_fx(p) + pt_fx(q);
                SY[l] = pt_fy(p) + pt_fy(q);
                SK[l] = (uint64_t(morton2(pt_x(q), pt_y(q))) << 16) | uint64_t(l);
                CL[l] = kNoClaim;
            } else {
                SK[l] = ~0ull;
            }
        }
        __syncthreads();
        // bitonic sort of SK[0, P): keys are distinct (the slot is in them)
        for (int k = 2; k <= P; k <<= 1)
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int i = tid; i < P; i += NT) {
                    const int ij = i ^ j;
                    if (ij > i) {
                        const uint64_t a = SK[i], b = SK[ij];
                        if ((a > b) == ((i & k) == 0)) {
                            SK[i] = b;
                            SK[ij] = a;
                        }
                    }
                }
                __syncthreads();
            }
        for (int k = tid; k < n; k += NT) RK[SK[k] & 0xFFFFu] = uint16_t(k);
        __syncthreads();
        if (n >= 2) {
            const uint64_t sid = pass_stream(uint64_t(c), pass, uint64_t(g));
            for (int r = 0; r < rounds; ++r) {
                uint32_t part[PER], pri[PER];
#pragma unroll
                for (int k = 0; k < PER; ++k) {
                    const int l = tid + k * NT;
                    part[k] = 0;
                    pri[k] = kNoClaim;
                    if (l < n) {
                        const uint64_t d = knn_draw(seed, sid, uint32_t(r), uint32_t(l));
                        part[k] = knn_partner(SK, RK, C, n, uint32_t(l), d);
                        pri[k] = (uint32_t(d) & 0xFFFFF000u) | uint32_t(l);
                        atomicMin(&CL[l], pri[k]);
                        atomicMin(&CL[part[k]], pri[k]);
                    }
                }
                __syncthreads();
#pragma unroll
                for (int k = 0; k < PER; ++k) {
                    const int i = tid + k * NT;
                    if (i < n && CL[i] == pri[k] && CL[part[k]] == pri[k]) {
                        const int j = int(part[k]);
                        const uint64_t ci = C[i], cj = C[j];
                        const int64_t d = swap_test(ci, cj, SX[i], SY[i], SX[j], SY[j]);
                        ++att;
                        if (d <= 0) {
                            C[i] = cj;
                            C[j] = ci;
                            ++acc;
                            gain += -2 * d;
                        }
                    }
                }
                __syncthreads();
                for (int l = tid; l < n; l += NT) CL[l] = kNoClaim;
                __syncthreads();
            }
        }
        for (int l = tid; l < n; l += NT) {
            const uint32_t s = permuted ? feistel_perm(uint32_t(g0 + l), uint32_t(w), pkey) : uint32_t(l);
            row[s] = C[l];
        }
        __syncthreads();
    }
    flush_stats<NT>(st, att, acc, gain);
    if (tid == 0) {
        __threadfence();
        const uint32_t done = atomicAdd(&queue[1], 1u);
        if (done == gridDim.x - 1) {
            atomicExch(&queue[0], 0u);
            atomicExch(&queue[1], 0u);
            atomicAdd(&queue[2], 1u);
        }
    }
}

struct TileVariant {
    int cap, threads;
    void (*fn)(uint64_t *, int, int64_t, const int64_t *, const int32_t *, const int32_t *, int, uint64_t, int,
               uint64_t, uint32_t *, amh_refine_stats *);
};
const TileVariant kTiles[] = {
    {1024, 256, k_refine_pass<1024, 256>},  {2048, 256, k_refine_pass<2048, 256>},
    {4096, 256, k_refine_pass<4096, 256>},  {4096, 512, k_refine_pass<4096, 512>},
    {8192, 512, k_refine_pass<8192, 512>},  {8192, 1024, k_refine_pass<8192, 1024>},
};
constexpr int kNumTiles = int(sizeof(kTiles) / sizeof(kTiles[0]));
int g_tile_grid[kNumTiles] = {0};
// kNN-guided variants of the first four tiles (slot indices fit 12 bits)
const TileVariant kKnnTiles[] = {
    {1024, 256, k_refine_knn<1024, 256>},
    {2048, 256, k_refine_knn<2048, 256>},
    {4096, 256, k_refine_knn<4096, 256>},
    {4096, 512, k_refine_knn<4096, 512>},
};
constexpr int kNumKnnTiles = int(sizeof(kKnnTiles) / sizeof(kKnnTiles[0]));
int g_knn_grid[kNumKnnTiles] = {0};

// Workgroups that are resident at once: the persistent grid of a variant.
int resident_grid(const TileVariant &v, int &cached) {
    if (cached == 0) {
        int dev = 0, ncu = 0, occ = 0;
        if (hipGetDevice(&dev) != hipSuccess) return 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void *>(v.fn), v.threads, 0) !=
            hipSuccess)
            return 0;
        cached = std::max(1, ncu * std::max(occ, 1));
    }
    return cached;
}
int tile_grid(int tile) { return resident_grid(kTiles[tile], g_tile_grid[tile]); }

}  // namespace

extern "C" int amh_chain_cost(const uint64_t *pts_dev, int32_t nkeys, int64_t wtot, const int64_t *chain_off_dev,
                              const int32_t *chain_w_dev, int32_t nchains, uint64_t *cost_dev, void *stream) {
    if (nkeys <= 0 || wtot < 0 || nchains <= 0 || !pts_dev || !chain_off_dev || !chain_w_dev || !cost_dev) {
        set_error("amh_chain_cost: invalid argument");
        return AMH_E_ARG;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = hipMemsetAsync(cost_dev, 0, sizeof(uint64_t) * 2 * size_t(nchains), s);
    if (e != hipSuccess) return check_hip(e, "amh_chain_cost memset");
    if (wtot == 0) return AMH_OK;
    const int64_t blocks = (wtot + kCostThreads - 1) / kCostThreads;
    hipLaunchKernelGGL(k_chain_cost, dim3(uint32_t(blocks)), dim3(kCostThreads), 0, s, pts_dev, nkeys, wtot,
                       chain_off_dev, chain_w_dev, nchains, reinterpret_cast<unsigned long long *>(cost_dev));
    AMH_LAUNCH_CHECK("amh_chain_cost");
    return AMH_OK;
}

extern "C" int amh_refine_chain(uint64_t *pts_dev, int32_t nkeys, int64_t wtot, int64_t chain_off, int32_t chain_w,
                                int32_t y, uint64_t rounds, uint64_t seed, uint64_t stream_id,
                                amh_refine_stats *stats_dev, void *stream) {
    if (!pts_dev || !stats_dev || nkeys <= 0 || chain_w < 0 || chain_off < 0 || chain_off + chain_w > wtot ||
        y < 0 || y >= nkeys) {
        set_error("amh_refine_chain: invalid argument (keys %d, off %lld, w %d, wtot %lld, y %d)", nkeys,
                  (long long)chain_off, chain_w, (long long)wtot, y);
        return AMH_E_ARG;
    }
    // One key frame: prev == next == the row itself, swaps never change the
    // cost; nothing to refine.  Chains of width < 2 have no pairs.
    if (nkeys < 2 || chain_w < 2) return AMH_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (rounds > 1 && chain_w <= kVisitCap) {
        // the whole visit in one launch
        amh::launch_timed("k_refine_visit", k_refine_visit, dim3(1), dim3(kVisitThreads), 0, s, pts_dev, int(nkeys),
                          wtot, chain_off, chain_w, y, seed, stream_id, rounds, stats_dev);
        AMH_LAUNCH_CHECK("amh_refine_chain");
        return AMH_OK;
    }
    for (uint64_t r = 0; r < rounds; ++r) {
        const RoundPar rp = round_params(seed, stream_id, r, uint32_t(chain_w));
        if (rp.npairs == 0) continue;
        const uint32_t blocks = (rp.npairs + kRoundThreads - 1) / kRoundThreads;
        hipLaunchKernelGGL(k_refine_round, dim3(blocks), dim3(kRoundThreads), 0, s, pts_dev, nkeys, wtot, chain_off,
                           chain_w, y, rp, stats_dev);
    }
    AMH_LAUNCH_CHECK("amh_refine_chain");
    return AMH_OK;
}

extern "C" int amh_refine_tiles(int32_t *count, int32_t *capacity, int32_t *threads, int32_t max_count) {
    if (count) *count = kNumTiles;
    for (int i = 0; i < kNumTiles && i < max_count; ++i) {
        if (capacity) capacity[i] = kTiles[i].cap;
        if (threads) threads[i] = kTiles[i].threads;
    }
    return AMH_OK;
}

extern "C" int amh_refine_items(const int32_t *chain_w, int32_t nchains, int32_t tile, int32_t *items,
                                int32_t max_items, int32_t *nitems) {
    if (!chain_w || !nitems || tile < 0 || tile >= kNumTiles || nchains < 0) {
        set_error("amh_refine_items: invalid argument");
        return AMH_E_ARG;
    }
    const int cap = kTiles[tile].cap;
    struct It { int c, g, n; };
    std::vector<It> v;
    for (int c = 0; c < nchains; ++c) {
        const int w = chain_w[c];
        if (w < 2) continue;
        const int ng = (w + cap - 1) / cap, gs = (w + ng - 1) / ng;
        for (int g = 0; g < ng; ++g) v.push_back({c, g, std::min(gs, w - g * gs)});
    }
    // Longest work first, so the persistent grid finishes together.
    std::stable_sort(v.begin(), v.end(), [](const It &a, const It &b) { return a.n > b.n; });
    *nitems = int32_t(v.size());
    if (items) {
        if (int32_t(v.size()) > max_items) {
            set_error("amh_refine_items: %d items exceed the buffer (%d)", int(v.size()), max_items);
            return AMH_E_ARG;
        }
        for (size_t i = 0; i < v.size(); ++i) {
            items[2 * i] = v[i].c;
            items[2 * i + 1] = v[i].g;
        }
    }
    return AMH_OK;
}

extern "C" int amh_refine_batched(uint64_t *pts_dev, int32_t nkeys, int64_t wtot, const int64_t *chain_off_dev,
                                  const int32_t *chain_w_dev, const int32_t *items_dev, int32_t nitems, int32_t tile,
                                  uint64_t pass, int32_t rounds, uint64_t seed, uint32_t *queue_dev,
                                  amh_refine_stats *stats_dev, void *stream) {
    if (!pts_dev || !chain_off_dev || !chain_w_dev || !queue_dev || !stats_dev || tile < 0 || tile >= kNumTiles ||
        nkeys <= 0 || nitems < 0 || rounds < 0) {
        set_error("amh_refine_batched: invalid argument");
        return AMH_E_ARG;
    }
    if (nkeys < 2 || nitems == 0 || rounds == 0) return AMH_OK;
    if (!items_dev) {
        set_error("amh_refine_batched: items_dev is NULL");
        return AMH_E_ARG;
    }
    const int grid = std::min(tile_grid(tile), int(nitems));
    if (grid <= 0) return check_hip(hipGetLastError(), "amh_refine_batched occupancy");
    hipStream_t s = static_cast<hipStream_t>(stream);
    amh::launch_timed("k_refine_pass", kTiles[tile].fn, dim3(grid), dim3(kTiles[tile].threads), 0, s, pts_dev, nkeys,
                      wtot, chain_off_dev, chain_w_dev, items_dev, int(nitems), pass, int(rounds), seed, queue_dev,
                      stats_dev);
    AMH_LAUNCH_CHECK("amh_refine_batched");
    return AMH_OK;
}

extern "C" int amh_refine_knn(uint64_t *pts_dev, int32_t nkeys, int64_t wtot, const int64_t *chain_off_dev,
                              const int32_t *chain_w_dev, const int32_t *items_dev, int32_t nitems, int32_t tile,
                              uint64_t pass, int32_t rounds, uint64_t seed, uint32_t *queue_dev,
                              amh_refine_stats *stats_dev, void *stream) {
    if (!pts_dev || !chain_off_dev || !chain_w_dev || !queue_dev || !stats_dev || tile < 0 || tile >= kNumKnnTiles ||
        nkeys <= 0 || nitems < 0 || rounds < 0) {
        set_error("amh_refine_knn: invalid argument (tile %d of %d kNN variants)", tile, kNumKnnTiles);
        return AMH_E_ARG;
    }
    if (nkeys < 2 || nitems == 0 || rounds == 0) return AMH_OK;
    if (!items_dev) {
        set_error("amh_refine_knn: items_dev is NULL");
        return AMH_E_ARG;
    }
    const int grid = std::min(resident_grid(kKnnTiles[tile], g_knn_grid[tile]), int(nitems));
    if (grid <= 0) return check_hip(hipGetLastError(), "amh_refine_knn occupancy");
    hipStream_t s = static_cast<hipStream_t>(stream);
    amh::launch_timed("k_refine_knn", kKnnTiles[tile].fn, dim3(grid), dim3(kKnnTiles[tile].threads), 0, s, pts_dev,
                      nkeys, wtot, chain_off_dev, chain_w_dev, items_dev, int(nitems), pass, int(rounds), seed,
                      queue_dev, stats_dev);
    AMH_LAUNCH_CHECK("amh_refine_knn");
    return AMH_OK;
}

// ------------------------------------------------------- graph replay
// A hipGraph of `npasses` batched launches that take their pass number from
// the queue's counter: one replay = npasses consecutive, distinct passes.
struct amh_graph_s {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" int amh_refine_graph_create(void **graph, uint64_t *pts_dev, int32_t nkeys, int64_t wtot,
                                       const int64_t *chain_off_dev, const int32_t *chain_w_dev,
                                       const int32_t *items_dev, int32_t nitems, int32_t tile, int32_t npasses,
                                       int32_t rounds, uint64_t seed, uint32_t *queue_dev,
                                       amh_refine_stats *stats_dev, void *stream) {
    if (!graph || !stream || npasses <= 0 || !pts_dev || !chain_off_dev || !chain_w_dev || !items_dev ||
        !queue_dev || !stats_dev || tile < 0 || tile >= kNumTiles || nkeys < 2 || nitems <= 0 || rounds <= 0) {
        set_error("amh_refine_graph_create: invalid argument (a created stream is required)");
        return AMH_E_ARG;
    }
    const int grid = std::min(tile_grid(tile), int(nitems));
    if (grid <= 0) return check_hip(hipGetLastError(), "amh_refine_graph_create occupancy");
    hipStream_t s = static_cast<hipStream_t>(stream);
    amh_graph_s *g = new amh_graph_s;
    int r = check_hip(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    if (!r) {
        for (int p = 0; p < npasses; ++p)
            hipLaunchKernelGGL(kTiles[tile].fn, dim3(grid), dim3(kTiles[tile].threads), 0, s, pts_dev, nkeys, wtot,
                               chain_off_dev, chain_w_dev, items_dev, int(nitems), kPassFromQueue, int(rounds), seed,
                               queue_dev, stats_dev);
        const hipError_t le = hipGetLastError();
        const hipError_t ee = hipStreamEndCapture(s, &g->graph);
        r = check_hip(le != hipSuccess ? le : ee, "graph capture");
    }
    if (!r) r = check_hip(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    if (r) {
        if (g->graph) (void)hipGraphDestroy(g->graph);
        delete g;
        return r;
    }
    *graph = g;
    return AMH_OK;
}

extern "C" int amh_graph_launch(void *graph, void *stream) {
    amh_graph_s *g = static_cast<amh_graph_s *>(graph);
    if (!g || !g->exec) {
        set_error("amh_graph_launch: NULL graph");
        return AMH_E_ARG;
    }
    amh::ProfScope ps_("refine_graph", static_cast<hipStream_t>(stream));
    return check_hip(hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream)), "hipGraphLaunch");
}

extern "C" int amh_graph_destroy(void *graph) {
    amh_graph_s *g = static_cast<amh_graph_s *>(graph);
    if (!g) return AMH_OK;
    (void)hipDeviceSynchronize();  // replays may still be queued on the caller's stream
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return AMH_OK;
}
