// lekf_regional.hip -- regional (SCALE-style) obs index and capped selection (gfx950).
//
// Obs index (scale/letkf/letkf_obs.f90:740-830, one subdomain): obs go to
// cell (ceiling(ri-0.5), ceiling(rj-0.5)) clamped to the grid, keeping input
// order within a cell (stable sort of unique keys cell << 32 | index).
//
// Selection (scale/letkf/letkf_tools.f90:853-1178): the rectangle of cells
// floor(ri -/+ dlon_zero) .. ceiling(...) is scanned row by row; a candidate
// passes when ndist = (sqrt(rdx^2+rdy^2)/sigma_h)^2 + (|dz|/sigma_v)^2 is not
// above the literal 13.33333333, with rloc = exp(-ndist/2) and
// rdiag = err^2/rloc.  With MAX_NOBS_PER_GRID = nobs_max > 0 every bucket
// (the reference's (element, obs type) pair) keeps its nobs_max smallest
// (rdiag, scan position) pairs -- exactly the set the reference's insertion
// sort keeps (strict < keeps earlier obs on ties).  Instead of inserting
// serially, a workgroup per point evaluates every candidate once into a
// scratch list in scan order, finds each bucket's threshold pair with
// radix-digit histogram passes over the list's 64-bit keys (10-bit digits in
// LDS), resolves the last digit bin by sorting at most LEKF_REG_LIST entries,
// and emits the kept obs bucket by bucket in scan order.  Without a cap the
// accepted obs are emitted in scan order.
#include "lekf_dev.hpp"
#include "lekf_regional.hpp"
#include "lekf_tune.hpp"

// Every FP64 operation in this file is rounded as written: no FMA contraction
// (the reference's results are reproduced operation by operation).
#pragma clang fp contract(off)

namespace {

constexpr double kNdMax = 13.33333333;  // scale/letkf/letkf_obs.f90:30

__device__ __forceinline__ unsigned long long lane_lt(int lane)
{
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

constexpr int kSelT = 256;  // threads per workgroup of the selection (4 waves)
constexpr int kSelW = kSelT / 64;
constexpr int kSelU = 1;    // candidates / list entries per thread per round

struct SelLds {
    int hist[LEKF_REG_NB][1024];
    unsigned long long lbits[LEKF_REG_NB][LEKF_REG_LIST];
    int lpos[LEKF_REG_NB][LEKF_REG_LIST];
    int seg_beg[LEKF_REG_SEG];
    int seg_pre[LEKF_REG_SEG + 1];
    int cnt[LEKF_REG_NB];
    int need[LEKF_REG_NB];
    int bincnt[LEKF_REG_NB];
    int mode[LEKF_REG_NB];  // 0 keep all, 1 refine, 2 collect, 3 tie scan
    int shift[LEKF_REG_NB];
    int lcnt[LEKF_REG_NB];
    int thr_pos[LEKF_REG_NB];
    int run[LEKF_REG_NB];
    int base[LEKF_REG_NB + 1];
    unsigned long long pref[LEKF_REG_NB];
    unsigned long long thr_bits[LEKF_REG_NB];
    int wcnt[kSelW][LEKF_REG_NB];
    int ecnt[kSelW][kSelU][LEKF_REG_NB];
    int any;
};

struct Cand {
    bool ok;
    int o, b;
    double rho, rd;
    unsigned long long bits;
};

struct ObsRec {
    double ri, rj, z, err;
    int b;
};

__device__ __forceinline__ ObsRec load_obs(const RegionalArgs& A, int o)
{
    ObsRec r;
    r.ri = A.ori[o];
    r.rj = A.orj[o];
    r.z = A.oz[o];
    r.err = A.oerr[o];
    r.b = A.obkt[o];
    return r;
}

// The reference's test and weights (scale/letkf/letkf_tools.f90:978-1037),
// each operation rounded as written (no contraction), so the cutoff decision
// and rdiag are bit-identical to the reference's.
__device__ __forceinline__ Cand eval_rec(const RegionalArgs& A, const ObsRec& r, int o, double ri,
                                         double rj, double rz)
{
    Cand c;
    const double rdx = rn_mul(rn_sub(ri, r.ri), A.dx);
    const double rdy = rn_mul(rn_sub(rj, r.rj), A.dy);
    const double nd_h = rn_div(__dsqrt_rn(rn_add(rn_mul(rdx, rdx), rn_mul(rdy, rdy))),
                                  A.sigma_h);
    const double nd_v = rn_div(fabs(rn_sub(r.z, rz)), A.sigma_v);
    const double nd = rn_add(rn_mul(nd_h, nd_h), rn_mul(nd_v, nd_v));
    c.o = o;
    c.ok = !(nd > kNdMax);
    c.b = r.b;
    c.rho = 0.0;
    c.rd = 0.0;
    c.bits = 0;
    if (c.ok) {
        c.rho = exp(-0.5 * nd);
        c.rd = rn_div(rn_mul(r.err, r.err), c.rho);
        c.bits = (unsigned long long)__double_as_longlong(c.rd);
    }
    return c;
}

// candidate q of the point's scan (row segments in LDS)
__device__ __forceinline__ int cand_obs(const SelLds& S, int nseg, int q, int& seg)
{
    while (seg + 1 < nseg && S.seg_pre[seg + 1] <= q) ++seg;
    return S.seg_beg[seg] + (q - S.seg_pre[seg]);
}

// Digits of a positive double's bit pattern (monotone in the value): shifts
// 53, 43, 33, 23, 13, 3 take 10 bits (bit 63 is the zero sign bit), shift 0
// the last 3 bits.
__device__ __forceinline__ int digit_width(int shift) { return shift == 0 ? 3 : 10; }

__device__ __forceinline__ int digit(unsigned long long bits, int shift)
{
    return (int)((bits >> shift) & ((1ull << digit_width(shift)) - 1ull));
}

// the bits at and above shift + width(shift) are already fixed in pref;
// shift < 0 means every bit is fixed
__device__ __forceinline__ bool prefix_match(unsigned long long bits, unsigned long long pref,
                                             int shift)
{
    if (shift < 0) return bits == pref;
    const int fixed_from = shift + digit_width(shift);
    if (fixed_from >= 63) return true;
    return (bits >> fixed_from) == (pref >> fixed_from);
}

__device__ __forceinline__ int next_shift(int shift)
{
    return shift == 3 ? 0 : (shift == 0 ? -1 : shift - 10);
}

// Block-ordered append: every thread with flag set gets a distinct position
// in block (thread) order after the S.run[slot] entries already placed; the
// per-slot running count advances by the block's total.  Uniform call.
__device__ __forceinline__ int block_append(SelLds& S, bool flag, int slot, int tid)
{
    const int lane = tid & 63, wv = tid >> 6;
    const unsigned long long m = __ballot(flag);
    if (lane == 0) S.wcnt[wv][slot] = __popcll(m);
    __syncthreads();
    int off = S.run[slot];
    for (int w = 0; w < wv; ++w) off += S.wcnt[w][slot];
    const int pos = off + __popcll(m & lane_lt(lane));
    __syncthreads();
    if (tid == 0) {
        int t = 0;
        for (int w = 0; w < kSelW; ++w) t += S.wcnt[w][slot];
        S.run[slot] += t;
    }
    __syncthreads();
    return pos;
}

// Keep, per bucket, the cap entries with the smallest (key, list position) of
// a list in scan order (key = rdiag bits, ob = obs | bucket << 30, rho) and
// write them bucket by bucket in list order to out_idx / out_rho; returns the
// count.  Each over-full bucket's threshold pair is found by digit histograms
// over the list.  Uniform call by all kSelT threads.
__device__ int cap_select(SelLds& S, int tid, const unsigned long long* skey, const int* sob,
                          const double* srho, int nacc, int nb, int cap, int* out_idx,
                          double* out_rho)
{
    for (int t = tid; t < LEKF_REG_NB; t += kSelT) {
        S.cnt[t] = 0;
        S.lcnt[t] = 0;
        S.mode[t] = 0;
        S.run[t] = 0;
    }
    for (int t = tid; t < LEKF_REG_NB * 1024; t += kSelT) (&S.hist[0][0])[t] = 0;
    __syncthreads();
    for (int i = tid; i < nacc; i += kSelT) {
        const int t = sob[i] >> 30;
        atomicAdd(&S.cnt[t], 1);
        atomicAdd(&S.hist[t][digit(skey[i], 53)], 1);
    }
    __syncthreads();
    // boundary digit of every bucket over its cap
    if (tid < nb) {
        const int t = tid;
        if (S.cnt[t] > cap) {
            int need = cap, d = 0, acc = 0;
            for (; d < 1024; ++d) {
                if (acc + S.hist[t][d] >= need) break;
                acc += S.hist[t][d];
            }
            S.need[t] = need - acc;
            S.pref[t] = (unsigned long long)d << 53;
            S.bincnt[t] = S.hist[t][d];
            S.shift[t] = 43;
            S.mode[t] = S.bincnt[t] <= LEKF_REG_LIST ? 2 : 1;
        }
    }
    __syncthreads();
    // refine while a boundary bin is too large to sort
    for (int pass = 0; pass < 8; ++pass) {
        if (tid == 0) {
            int a = 0;
            for (int t = 0; t < nb; ++t) a |= (S.mode[t] == 1);
            S.any = a;
        }
        __syncthreads();
        if (!S.any) break;
        for (int t = 0; t < nb; ++t)
            if (S.mode[t] == 1)
                for (int u = tid; u < 1024; u += kSelT) S.hist[t][u] = 0;
        __syncthreads();
        for (int i = tid; i < nacc; i += kSelT) {
            const unsigned long long bits = skey[i];
            const int t = sob[i] >> 30;
            if (S.mode[t] == 1 && prefix_match(bits, S.pref[t], S.shift[t]))
                atomicAdd(&S.hist[t][digit(bits, S.shift[t])], 1);
        }
        __syncthreads();
        if (tid < nb && S.mode[tid] == 1) {
            const int t = tid;
            const int sh = S.shift[t];
            int need = S.need[t], d = 0, acc = 0;
            for (; d < 1024; ++d) {
                if (acc + S.hist[t][d] >= need) break;
                acc += S.hist[t][d];
            }
            S.need[t] = need - acc;
            S.pref[t] |= (unsigned long long)d << sh;
            S.bincnt[t] = S.hist[t][d];
            const int ns = next_shift(sh);
            S.shift[t] = ns;
            if (S.bincnt[t] <= LEKF_REG_LIST)
                S.mode[t] = 2;
            else if (ns < 0)
                S.mode[t] = 3;  // every key bit fixed: equal keys, break ties by position
        }
        __syncthreads();
    }
    // collect the boundary bins; resolve exact ties in list (scan) order
    for (int t = tid; t < LEKF_REG_NB; t += kSelT) S.run[t] = 0;
    __syncthreads();
    for (int i0 = 0; i0 < nacc; i0 += kSelT) {
        const int i = i0 + tid;
        bool in = i < nacc;
        const unsigned long long bits = in ? skey[i] : 0ull;
        const int t = in ? (sob[i] >> 30) : 0;
        if (in && S.mode[t] == 2 && prefix_match(bits, S.pref[t], S.shift[t])) {
            const int pos = atomicAdd(&S.lcnt[t], 1);
            if (pos < LEKF_REG_LIST) {
                S.lbits[t][pos] = bits;
                S.lpos[t][pos] = i;
            }
        }
        for (int u = 0; u < nb; ++u) {
            if (S.mode[u] != 3) continue;
            const bool tie = in && t == u && bits == S.pref[u];
            const int rank = block_append(S, tie, u, tid);
            if (tie && rank == S.need[u] - 1) {
                S.thr_bits[u] = bits;
                S.thr_pos[u] = i;
            }
        }
    }
    __syncthreads();
    // sort each collected bin by (bits, position); the need-th is the threshold
    for (int t = 0; t < nb; ++t) {
        if (S.mode[t] != 2) continue;
        const int n = S.lcnt[t] < LEKF_REG_LIST ? S.lcnt[t] : LEKF_REG_LIST;
        for (int ph = 0; ph < n; ++ph) {
            const int i = 2 * tid + (ph & 1);
            if (i + 1 < n) {
                const unsigned long long x = S.lbits[t][i], y = S.lbits[t][i + 1];
                const int px = S.lpos[t][i], py = S.lpos[t][i + 1];
                if (x > y || (x == y && px > py)) {
                    S.lbits[t][i] = y;
                    S.lbits[t][i + 1] = x;
                    S.lpos[t][i] = py;
                    S.lpos[t][i + 1] = px;
                }
            }
            __syncthreads();
        }
        if (tid == 0) {
            S.thr_bits[t] = S.lbits[t][S.need[t] - 1];
            S.thr_pos[t] = S.lpos[t][S.need[t] - 1];
        }
        __syncthreads();
    }
    if (tid == 0) {
        int acc = 0;
        for (int t = 0; t < nb; ++t) {
            S.base[t] = acc;
            acc += S.cnt[t] < cap ? S.cnt[t] : cap;
            S.run[t] = 0;
        }
        S.base[nb] = acc;
    }
    __syncthreads();
    // emit bucket by bucket in scan order (list position i = i0 + u * kSelT + tid)
    {
        const int lane = tid & 63, wv = tid >> 6;
        int runb[LEKF_REG_NB];
#pragma unroll
        for (int t = 0; t < LEKF_REG_NB; ++t) runb[t] = 0;
        for (int i0 = 0; i0 < nacc; i0 += kSelT * kSelU) {
            unsigned long long bits[kSelU];
            int ob[kSelU];
            bool keep[kSelU];
#pragma unroll
            for (int u = 0; u < kSelU; ++u) {
                const int i = i0 + u * kSelT + tid;
                const bool in = i < nacc;
                bits[u] = in ? skey[i] : 0ull;
                ob[u] = in ? sob[i] : 0;
                const int t = ob[u] >> 30;
                keep[u] = in && (S.mode[t] == 0 || bits[u] < S.thr_bits[t] ||
                                 (bits[u] == S.thr_bits[t] && i <= S.thr_pos[t]));
            }
            unsigned long long m[kSelU][LEKF_REG_NB];
#pragma unroll
            for (int u = 0; u < kSelU; ++u)
#pragma unroll
                for (int t = 0; t < LEKF_REG_NB; ++t) m[u][t] = __ballot(keep[u] && (ob[u] >> 30) == t);
            if (lane == 0)
#pragma unroll
                for (int u = 0; u < kSelU; ++u)
#pragma unroll
                    for (int t = 0; t < LEKF_REG_NB; ++t) S.ecnt[wv][u][t] = __popcll(m[u][t]);
            __syncthreads();
            int pos[kSelU];
#pragma unroll
            for (int u = 0; u < kSelU; ++u) pos[u] = -1;
#pragma unroll
            for (int u = 0; u < kSelU; ++u)
#pragma unroll
                for (int t = 0; t < LEKF_REG_NB; ++t) {
                    int before = 0, tot = 0;
                    for (int w = 0; w < kSelW; ++w) {
                        const int cw = S.ecnt[w][u][t];
                        before += w < wv ? cw : 0;
                        tot += cw;
                    }
                    if (keep[u] && (ob[u] >> 30) == t)
                        pos[u] = runb[t] + before + __popcll(m[u][t] & lane_lt(lane));
                    runb[t] += tot;
                }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < kSelU; ++u) {
                const int t = ob[u] >> 30;
                if (keep[u] && pos[u] < S.base[t + 1] - S.base[t]) {
                    const long w = S.base[t] + pos[u];
                    out_idx[w] = ob[u] & 0x3fffffff;
                    out_rho[w] = srho[i0 + u * kSelT + tid];
                }
            }
        }
    }
    const int kept = S.base[nb];
    __syncthreads();
    return kept;
}

// Persistent selection: workgroup g handles points g, g + gridDim.x, ...
// Each candidate of a point is evaluated once.  Without a cap the accepted
// candidates are written in scan order.  With a cap they go to the
// workgroup's scratch slot in scan order (rdiag bits, obs | bucket << 30,
// rho); each over-full bucket's threshold pair (key, list position) is
// found by digit histograms over the list, and the kept entries are written
// bucket by bucket in scan order.
__global__ __launch_bounds__(kSelT) void reg_select(RegionalArgs A, long npt)
{
    __shared__ SelLds S;
    const int tid = threadIdx.x;
    const int nij = A.nx * A.ny;
    const int nb = A.nbkt;
    const int cap = A.nobs_max;
    unsigned long long* skey = A.s_key + (size_t)blockIdx.x * A.slot_cap;
    int* sob = A.s_ob + (size_t)blockIdx.x * A.slot_cap;
    double* srho = A.s_rho + (size_t)blockIdx.x * A.slot_cap;

    for (long pi = blockIdx.x; pi < npt; pi += gridDim.x) {
        const long b = A.pt_begin + pi;  // global point = lev * nij + ij
        const int lev = (int)(b / nij), ij = (int)(b % nij);
        const double ri = (double)(ij % A.nx + 1), rj = (double)(ij / A.nx + 1), rz = A.zlev[lev];
        int imin = (int)floor(ri - A.dlon_zero), imax = (int)ceil(ri + A.dlon_zero);
        int jmin = (int)floor(rj - A.dlat_zero), jmax = (int)ceil(rj + A.dlat_zero);
        if (imin < 1) imin = 1;
        if (imax > A.nx) imax = A.nx;
        if (jmin < 1) jmin = 1;
        if (jmax > A.ny) jmax = A.ny;
        const int nseg = (imin > imax || jmin > jmax) ? 0 : jmax - jmin + 1;
        if (tid == 0) {
            int acc = 0;
            for (int s = 0; s < nseg; ++s) {
                const int j = jmin + s;  // 1-based row
                // cellcum = inclusive cumulative count through cell (i, j), row-major
                const long rowbase = (long)(j - 1) * A.nx;
                const int beg = (imin > 1) ? A.cellcum[rowbase + imin - 2]
                                           : (rowbase > 0 ? A.cellcum[rowbase - 1] : 0);
                const int end = A.cellcum[rowbase + imax - 1];
                S.seg_beg[s] = beg;
                S.seg_pre[s] = acc;
                acc += end - beg;
            }
            S.seg_pre[nseg] = acc;
        }
        __syncthreads();
        const int ncand = S.seg_pre[nseg];
        const long obase = A.list_off[b] - A.list_base;  // relative to this chunk's list buffer

        // evaluate every candidate once: kSelU per thread per round, all loads
        // of a round issued first; scan position q = q0 + u * kSelT + tid
        int nacc = 0;
        {
            const int lane = tid & 63, wv = tid >> 6;
            int seg = 0;
            for (int q0 = 0; q0 < ncand; q0 += kSelT * kSelU) {
                int o[kSelU];
                ObsRec r[kSelU];
#pragma unroll
                for (int u = 0; u < kSelU; ++u) {
                    const int q = q0 + u * kSelT + tid;
                    o[u] = q < ncand ? cand_obs(S, nseg, q, seg) : -1;
                    if (o[u] >= 0) r[u] = load_obs(A, o[u]);
                }
                Cand c[kSelU];
                unsigned long long m[kSelU];
#pragma unroll
                for (int u = 0; u < kSelU; ++u) {
                    c[u].ok = false;
                    c[u].b = 0;
                    if (o[u] >= 0) c[u] = eval_rec(A, r[u], o[u], ri, rj, rz);
                    m[u] = __ballot(c[u].ok);
                }
                if (lane == 0)
#pragma unroll
                    for (int u = 0; u < kSelU; ++u) S.wcnt[wv][u] = __popcll(m[u]);
                __syncthreads();
                int pos[kSelU];
#pragma unroll
                for (int u = 0; u < kSelU; ++u) {
                    int before = 0, tot = 0;
                    for (int w = 0; w < kSelW; ++w) {
                        const int cw = S.wcnt[w][u];
                        before += w < wv ? cw : 0;
                        tot += cw;
                    }
                    pos[u] = nacc + before + __popcll(m[u] & lane_lt(lane));
                    nacc += tot;
                }
                __syncthreads();
#pragma unroll
                for (int u = 0; u < kSelU; ++u) {
                    if (!c[u].ok) continue;
                    if (cap <= 0) {
                        A.loc_idx[obase + pos[u]] = c[u].o;
                        A.loc_rho[obase + pos[u]] = c[u].rho;
                    } else {
                        skey[pos[u]] = c[u].bits;
                        sob[pos[u]] = c[u].o | (c[u].b << 30);
                        srho[pos[u]] = c[u].rho;
                    }
                }
            }
        }
        __syncthreads();
        if (cap <= 0) {
            if (tid == 0) {
                A.loc_off[b] = obase;
                A.loc_cnt[b] = nacc;
            }
            __syncthreads();
            continue;
        }
        const int kept = cap_select(S, tid, skey, sob, srho, nacc, nb, cap, A.loc_idx + obase,
                                    A.loc_rho + obase);
        if (tid == 0) {
            A.loc_off[b] = obase;
            A.loc_cnt[b] = kept;
        }
        __syncthreads();
    }
}

// Nearest-first cap of per-point lists (persistent, one point per workgroup).
__global__ __launch_bounds__(kSelT) void list_cap(ListCapArgs A)
{
    __shared__ SelLds S;
    const int tid = threadIdx.x;
    unsigned long long* skey = A.s_key + (size_t)blockIdx.x * A.slot_cap;
    int* sob = A.s_ob + (size_t)blockIdx.x * A.slot_cap;
    double* srho = A.s_rho + (size_t)blockIdx.x * A.slot_cap;
    const long npt = (long)A.ncols * A.nlev;
    for (long pi = blockIdx.x; pi < npt; pi += gridDim.x) {
        const long b = (pi / A.ncols) * (long)A.nij + A.col_begin + pi % A.ncols;
        const long off = A.in_off[b];
        const int n = A.in_cnt[b];
        const long ob = b * (long)A.cap;
        int kept = n;
        if (n <= A.cap) {
            for (int i = tid; i < n; i += kSelT) {
                A.out_idx[ob + i] = A.in_idx[off + i];
                A.out_rho[ob + i] = A.in_rho[off + i];
            }
        } else {
            for (int i = tid; i < n; i += kSelT) {
                const int o = A.in_idx[off + i];
                const double rho = A.in_rho[off + i];
                skey[i] = (unsigned long long)__double_as_longlong(rn_div(A.rdiag[o], rho));
                sob[i] = o;
                srho[i] = rho;
            }
            __syncthreads();
            kept = cap_select(S, tid, skey, sob, srho, n, 1, A.cap, A.out_idx + ob, A.out_rho + ob);
        }
        if (tid == 0) {
            A.out_off[b] = ob;
            A.out_cnt[b] = kept;
        }
        __syncthreads();
    }
}

// per point: candidate count of the rectangle -> list capacity
__global__ void reg_capacity(RegionalArgs A, int* capn, int* maxcand)
{
    const long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long npt = (long)A.nx * A.ny * A.nlev;
    if (b >= npt) return;
    const int nij = A.nx * A.ny;
    const int ij = (int)(b % nij);
    const double ri = (double)(ij % A.nx + 1), rj = (double)(ij / A.nx + 1);
    int imin = (int)floor(ri - A.dlon_zero), imax = (int)ceil(ri + A.dlon_zero);
    int jmin = (int)floor(rj - A.dlat_zero), jmax = (int)ceil(rj + A.dlat_zero);
    if (imin < 1) imin = 1;
    if (imax > A.nx) imax = A.nx;
    if (jmin < 1) jmin = 1;
    if (jmax > A.ny) jmax = A.ny;
    long n = 0;
    for (int j = jmin; j <= jmax && imin <= imax; ++j) {
        const long rowbase = (long)(j - 1) * A.nx;
        const int beg = (imin > 1) ? A.cellcum[rowbase + imin - 2] : (rowbase > 0 ? A.cellcum[rowbase - 1] : 0);
        n += A.cellcum[rowbase + imax - 1] - beg;
    }
    atomicMax(maxcand, (int)n);
    if (A.nobs_max > 0) {
        const long c = (long)A.nobs_max * A.nbkt;
        n = n < c ? n : c;
    }
    capn[b] = (int)n;
}

// obs keys: cell (ceiling(ri-0.5), ceiling(rj-0.5)) clamped; qc-rejected obs sort last
__global__ void reg_keys(RegObsArgs A)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= A.nobs) return;
    unsigned long long key = ~0ull;
    if (!A.qc || A.qc[n] == A.qc_keep) {
        int i = (int)ceil(A.ri[n] - 0.5), j = (int)ceil(A.rj[n] - 0.5);
        i = i < 1 ? 1 : (i > A.nx ? A.nx : i);
        j = j < 1 ? 1 : (j > A.ny ? A.ny : j);
        const unsigned cell = (unsigned)((i - 1) + A.nx * (j - 1));
        key = ((unsigned long long)cell << 32) | (unsigned)n;
        atomicAdd(&A.cell_cnt[cell], 1);
    }
    A.keys[n] = key;
}

__global__ void reg_scatter(RegObsArgs A, int nbinned)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nbinned) return;
    const int n = (int)(A.keys[m] & 0xffffffffull);
    A.perm[m] = n;
    A.o_ri[m] = A.ri[n];
    A.o_rj[m] = A.rj[n];
    A.o_z[m] = A.z[n];
    A.o_err[m] = A.err[n];
    A.o_rdiag[m] = A.err[n] * A.err[n];
    A.o_dep[m] = A.dep[n];
    A.o_bkt[m] = A.bkt[n];
    for (int i = 0; i < A.ne; ++i) {
        const double v = A.hdxf[(size_t)n * A.h_so + (size_t)i * A.h_sm];
        A.o_hdxf[(size_t)m * A.ne + i] = v;
        if (A.o_hdxf32) A.o_hdxf32[(size_t)m * A.ne + i] = (float)v;
    }
}

__global__ __launch_bounds__(1024) void scan_inclusive_int(const int* in, int* out, int n)
{
    __shared__ int s[1024];
    int carry = 0;
    for (int base = 0; base < n; base += 1024) {
        const int i = base + threadIdx.x;
        const int v = i < n ? in[i] : 0;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            const int t = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
            __syncthreads();
            s[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < n) out[i] = carry + s[threadIdx.x];
        const int tot = s[1023];
        __syncthreads();
        carry += tot;
    }
}

__global__ __launch_bounds__(1024) void scan_exclusive_long(const int* in, long* out, long n)
{
    __shared__ long s[1024];
    long carry = 0;
    for (long base = 0; base < n; base += 1024) {
        const long i = base + threadIdx.x;
        const long v = i < n ? in[i] : 0;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            const long t = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
            __syncthreads();
            s[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < n) out[i] = carry + s[threadIdx.x] - v;
        const long tot = s[1023];
        __syncthreads();
        carry += tot;
    }
    if (threadIdx.x == 0) out[n] = carry;
}

}  // namespace

hipError_t lekf_reg_keys_launch(const RegObsArgs& A, hipStream_t st)
{
    if (A.nobs > 0)
        hipLaunchKernelGGL(reg_keys, dim3((A.nobs + 255) / 256), dim3(256), 0, st, A);
    return hipGetLastError();
}

hipError_t lekf_reg_scatter_launch(const RegObsArgs& A, int nbinned, int ncell, int* cellcum,
                                   hipStream_t st)
{
    if (nbinned > 0)
        hipLaunchKernelGGL(reg_scatter, dim3((nbinned + 255) / 256), dim3(256), 0, st, A, nbinned);
    hipLaunchKernelGGL(scan_inclusive_int, dim3(1), dim3(1024), 0, st, A.cell_cnt, cellcum, ncell);
    return hipGetLastError();
}

hipError_t lekf_reg_capacity_launch(const RegionalArgs& A, int* capn, long* cap_off, int* maxcand,
                                    hipStream_t st)
{
    const long npt = (long)A.nx * A.ny * A.nlev;
    hipLaunchKernelGGL(reg_capacity, dim3((unsigned)((npt + 255) / 256)), dim3(256), 0, st, A, capn,
                       maxcand);
    hipLaunchKernelGGL(scan_exclusive_long, dim3(1), dim3(1024), 0, st, capn, cap_off, npt);
    return hipGetLastError();
}

// workgroup slots of the persistent selection (LDS allows up to four per CU)
int lekf_reg_select_slots()
{
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    const int per = g_lekf_tune.sel_per_cu < 1 ? 1 : (g_lekf_tune.sel_per_cu > 4 ? 4 : g_lekf_tune.sel_per_cu);
    return per * (ncu > 0 ? ncu : 256);
}

int lekf_reg_select_slots_max()
{
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    return 4 * (ncu > 0 ? ncu : 256);
}

hipError_t lekf_list_cap_launch(const ListCapArgs& A, int nslots, hipStream_t st)
{
    const long npt = (long)A.ncols * A.nlev;
    if (npt <= 0) return hipSuccess;
    const long g = npt < nslots ? npt : nslots;
    hipLaunchKernelGGL(list_cap, dim3((unsigned)g), dim3(kSelT), 0, st, A);
    return hipGetLastError();
}

hipError_t lekf_reg_select_launch(const RegionalArgs& A, long npt, int nslots, hipStream_t st)
{
    if (npt <= 0) return hipSuccess;
    const long g = npt < nslots ? npt : nslots;
    hipLaunchKernelGGL(reg_select, dim3((unsigned)g), dim3(kSelT), 0, st, A, npt);
    return hipGetLastError();
}
