// ipc contact mode on gfx950: the primitive-distance barrier of rigid IPC over a list of
// point-triangle / edge-edge pairs of one scene (nexdynipc_amd/csrc/ipc_contact.h holds
// the per-pair mathematics, shared with the host).
//
//   k_ipc_filter    one lane per pair: the pair's check, segment key and whether it is
//                   within dhat;
//   k_ipc_seg_*     consecutive pairs of one unordered body pair form a segment: per
//                   chunk counts, one scan, per chunk writes give every segment's first
//                   pair and the in-order list of the pairs within dhat;
//   k_ipc_pairs     sixteen lanes per listed pair: world points and rotation Jacobians
//                   of the four vertices, distance type, E = kappa b(d) with its vertex
//                   gradient and Hessian (hyper-dual, split over lanes by derivative
//                   slot), the PSD projection (Jacobi rotations over lanes, in LDS), the
//                   fold into the two bodies' 12 DOFs; results pair-major in HBM;
//   k_ipc_reduce    one work-group per segment (persistent over segments), one lane per
//                   output element (E, 12 gradient, 144 Hessian entries), each summing
//                   its element over the segment's pairs within dhat in pair order;
//   k_ipc_total     the scene energy, the segments' energies summed in segment order;
//   k_ipc_advance   maxStep: one lane per pair, conservative advancement along the
//                   line-search path, then the least over the pairs.
//
// Every sum runs in a fixed order, so the results are bitwise those of the host
// evaluation of the same pairs. The per-pair work is FP64 VALU and LDS (most of it the
// 12 x 12 Jacobi eigen-decomposition of the projection: up to 66 dependent rotations
// per sweep) with a few hundred bytes of gathered input; the reduction streams 157
// doubles per pair.
#include "ipc_contact.h"
#include "ndipc_internal.h"

namespace nd {

constexpr int kIpcScanThreads = 1024;
constexpr int kIpcReduceThreads = 192;  // 157 outputs per segment, three wavefronts
constexpr int kIpcLanes = 16;           // lanes per pair
constexpr int kIpcPairsPerWg = 4;       // pairs per work-group (one wavefront)
constexpr int kIpcSlots = 5;            // second-derivative slots per lane: 16 x 5 >= 78

// One slot of an ipc::H12 hyper-dual: the value, the first derivatives d[i], d[j] and the
// second derivative h[(i, j)] of the slot's index pair (i <= j). Every operation of H12
// computes d[k] from d[k] alone and h[(i, j)] from h[(i, j)], d[i], d[j], so a lane that
// carries slot (i, j) through the same expression gets the same bits as the full H12.
struct HL {
    double v, di, dj, h;
};
__device__ __forceinline__ HL operator+(const HL& a, const HL& b) {
    return {a.v + b.v, a.di + b.di, a.dj + b.dj, a.h + b.h};
}
__device__ __forceinline__ HL operator-(const HL& a, const HL& b) {
    return {a.v - b.v, a.di - b.di, a.dj - b.dj, a.h - b.h};
}
__device__ __forceinline__ HL operator*(const HL& a, const HL& b) {
    return {a.v * b.v, a.di * b.v + a.v * b.di, a.dj * b.v + a.v * b.dj,
            (a.h * b.v + a.v * b.h) + (a.di * b.dj + a.dj * b.di)};
}
__device__ __forceinline__ HL operator/(const HL& a, const HL& b) {
    const double f = 1.0 / b.v, f1 = -f * f, f2 = -2.0 * f1 * f;
    const HL ib = {f, f1 * b.di, f1 * b.dj, f1 * b.h + f2 * b.di * b.dj};
    return a * ib;
}
__device__ __forceinline__ HL hl_var(double v, int var, int si, int sj) {
    return {v, si == var ? 1.0 : 0.0, sj == var ? 1.0 : 0.0, 0.0};
}
// (i, j) of packed upper-triangle index t (ipc::hix order)
__device__ __forceinline__ void slot_ij(int t, int& i, int& j) {
    i = 0;
    while (t >= ipc::kN - i) {
        t -= ipc::kN - i;
        ++i;
    }
    j = i + t;
}
// entry (r, c) of the 12 x 12 map from the two bodies' DOFs to the four vertices
// (ipc::fold_to_bodies' J)
__device__ __forceinline__ double ipc_jel(int r, int c, const int* own, const double* jth) {
    const int v = r / 3, a = r - 3 * v, o = 6 * own[v];
    if (c == o + a) return 1.0;
    if (c >= o + 3 && c < o + 6) return jth[9 * v + 3 * a + (c - o - 3)];
    return 0.0;
}

// Pass 1, one lane per pair: the pair's check, its segment key, and whether it is within
// dhat, from the same world points (ipc::vertex_and_jacobian) and the same squared
// distance value (the value part of the hyper-dual, whose division is a * (1 / b)) as
// pass 2, so both passes agree on every pair.
__global__ __launch_bounds__(256) void k_ipc_filter(Dev d, int b0, int nb, int np, const int* pairs, const double* x,
                                                    double dhat, double kappa, int value_only, unsigned long long* key,
                                                    int* act, int* bad, double* pe) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= np) return;
    const int* pr = pairs + 5 * i;
    const int kind = pr[0], a = pr[1], b = pr[2], ia = pr[3], ib = pr[4];
    bool ok = (kind == 0 || kind == 1) && a >= 0 && b >= 0 && a < nb && b < nb && a != b;
    const int ga = ok ? b0 + a : 0, gb = ok ? b0 + b : 0;
    ok = ok && d.nv[ga] > 0 && d.nv[gb] > 0 && ia >= 0 && ib >= 0;
    ok = ok && (kind == 0 ? (ia < d.nv[ga] && ib < d.nf[gb]) : (ia < d.ne[ga] && ib < d.ne[gb]));
    if (!ok) {
        key[i] = ~0ull;
        act[i] = 0;
        atomicMin(bad, i);
        return;
    }
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    key[i] = (static_cast<unsigned long long>(lo) << 32) | static_cast<unsigned long long>(hi);
    double xw[ipc::kN], jth[9];
    for (int l = 0; l < 4; ++l) {
        int g, vl;
        if (kind == 0) {
            g = l == 0 ? ga : gb;
            vl = l == 0 ? ia : d.faces[3 * (d.f_off[gb] + ib) + (l - 1)];
        } else {
            g = l < 2 ? ga : gb;
            vl = l < 2 ? d.edges[2 * (d.e_off[ga] + ia) + l] : d.edges[2 * (d.e_off[gb] + ib) + (l - 2)];
        }
        const double q[4] = {d.q[4 * g], d.q[4 * g + 1], d.q[4 * g + 2], d.q[4 * g + 3]};
        const double* r = d.verts + 3 * (d.v_off[g] + vl);
        const double rr[3] = {r[0], r[1], r[2]};
        ipc::vertex_and_jacobian(x + 6 * (g - b0), q, rr, xw + 3 * l, jth);
    }
    const int type = ipc::pair_type(kind, xw);
    ipc::P3<HL> v[4];
    for (int k = 0; k < 4; ++k)
        v[k] = {hl_var(xw[3 * k], 3 * k, 0, 0), hl_var(xw[3 * k + 1], 3 * k + 1, 0, 0),
                hl_var(xw[3 * k + 2], 3 * k + 2, 0, 0)};
    const HL d2 = ipc::pair_d2<HL>(kind, type, v);
    const double dist = sqrt(d2.v > 0.0 ? d2.v : 0.0);
    act[i] = dist < dhat ? 1 : 0;
    // value only (line searches): the pair's energy here, as pass 2 computes it
    if (value_only && dist < dhat) pe[i] = bar_v(dist, dhat, kappa);
}

// Pass 2 over the pairs within dhat (alist[0 .. *n_act), pair order), sixteen lanes per
// pair, four pairs per work-group of one wavefront, persistent over the list. Lanes 0-3
// map the four vertices; every lane carries up to five second-derivative slots of the
// squared distance; the PSD projection is ipc::psd_project with each rotation's column,
// row and eigenvector updates spread over lanes 0-11 (the rotation angle, the sweep test
// and the break are computed as there, in the same order); the fold takes one output
// entry per lane at a time. Control flow is uniform over the wavefront: a pair that is
// done (or a slot past the list's end) masks its lanes' work.
__global__ __launch_bounds__(64) void k_ipc_pairs(Dev d, int b0, const int* pairs, const double* x, double dhat,
                                                  double kappa, int psd, const int* alist, const int* n_act,
                                                  double* pe, double* pg, double* pH) {
    constexpr int N = ipc::kN;
    __shared__ double sA[kIpcPairsPerWg][N * N], sV[kIpcPairsPerWg][N * N], sP[kIpcPairsPerWg][N * N];
    __shared__ double sG[kIpcPairsPerWg][N], sJ[kIpcPairsPerWg][36], sX[kIpcPairsPerWg][N];
    __shared__ int sOwn[kIpcPairsPerWg][4], sFlag[kIpcPairsPerWg];
    const int grp = threadIdx.x / kIpcLanes, l = threadIdx.x % kIpcLanes;
    const int na = *n_act;
    for (int base = blockIdx.x * kIpcPairsPerWg; base < na; base += gridDim.x * kIpcPairsPerWg) {
    const int j = base + grp;
    const bool ok = j < na;
    const int i = ok ? alist[j] : 0;
    int kind = 0, a = 0, b = 0, ia = 0, ib = 0;
    if (ok) {
        const int* pr = pairs + 5 * i;
        kind = pr[0];
        a = pr[1];
        b = pr[2];
        ia = pr[3];
        ib = pr[4];
    }
    const int ga = b0 + a, gb = b0 + b;
    const int lo = a < b ? a : b;
    if (ok && l < 4) {
        int g, vl;
        if (kind == 0) {
            g = l == 0 ? ga : gb;
            vl = l == 0 ? ia : d.faces[3 * (d.f_off[gb] + ib) + (l - 1)];
        } else {
            g = l < 2 ? ga : gb;
            vl = l < 2 ? d.edges[2 * (d.e_off[ga] + ia) + l] : d.edges[2 * (d.e_off[gb] + ib) + (l - 2)];
        }
        const double q[4] = {d.q[4 * g], d.q[4 * g + 1], d.q[4 * g + 2], d.q[4 * g + 3]};
        const double* r = d.verts + 3 * (d.v_off[g] + vl);
        const double rr[3] = {r[0], r[1], r[2]};
        ipc::vertex_and_jacobian(x + 6 * (g - b0), q, rr, &sX[grp][3 * l], &sJ[grp][9 * l]);
        sOwn[grp][l] = (g - b0) == lo ? 0 : 1;
    }
    __syncthreads();
    double xw[N];
    for (int k = 0; k < N; ++k) xw[k] = ok ? sX[grp][k] : 0.0;
    const int type = ok ? ipc::pair_type(kind, xw) : 0;
    bool active = false;
    double dist = 0.0, E = 0.0, e1 = 0.0, e2 = 0.0;
#pragma unroll 1
    for (int s = 0; s < kIpcSlots; ++s) {
        const int t = l + kIpcLanes * s;
        int si, sj;
        slot_ij(t < ipc::kH ? t : 0, si, sj);
        ipc::P3<HL> v[4];
        for (int k = 0; k < 4; ++k)
            v[k] = {hl_var(xw[3 * k], 3 * k, si, sj), hl_var(xw[3 * k + 1], 3 * k + 1, si, sj),
                    hl_var(xw[3 * k + 2], 3 * k + 2, si, sj)};
        const HL d2 = ipc::pair_d2<HL>(kind == 1 ? 1 : 0, type, v);
        if (s == 0) {
            dist = sqrt(d2.v > 0.0 ? d2.v : 0.0);
            active = ok && dist < dhat;
            if (!active) break;  // no barrier term: the remaining slots are not needed
            E = bar_v(dist, dhat, kappa);
            if (dist > 0.0) {
                const double b1 = bar_g(dist, dhat, kappa), b2 = bar_h(dist, dhat, kappa);
                e1 = b1 / (2.0 * dist);
                e2 = (b2 - b1 / dist) / (4.0 * dist * dist);
            }
        }
        if (active && t < ipc::kH) {
            double hv = 0.0, gv = 0.0;
            if (dist > 0.0) {
                hv = e2 * d2.di * d2.dj + e1 * d2.h;
                gv = e1 * d2.di;
            }
            sA[grp][si * N + sj] = hv;
            sA[grp][sj * N + si] = hv;
            if (si == sj) sG[grp][si] = gv;
        }
    }
    if (active && l == 0) pe[i] = E;
    __syncthreads();
    const double* Hs = sA[grp];
    if (psd) {
        double* A = sA[grp];
        double* V = sV[grp];
        for (int idx = l; idx < N * N; idx += kIpcLanes) V[idx] = (idx / N == idx % N) ? 1.0 : 0.0;
        double tot = 0.0;
        if (active && l == 0)
            for (int k = 0; k < N * N; ++k) tot += A[k] * A[k];
        for (int sweep = 0; sweep < 50; ++sweep) {
            if (l == 0) {
                int done = 1;
                if (active) {
                    double off = 0.0;
                    for (int r = 0; r < N; ++r)
                        for (int c = r + 1; c < N; ++c) off += A[r * N + c] * A[r * N + c];
                    done = (off > 1e-30 * tot) ? 0 : 1;
                }
                sFlag[grp] = done;
            }
            __syncthreads();
            const bool gdone = sFlag[grp] != 0;
            const bool all = sFlag[0] && sFlag[1] && sFlag[2] && sFlag[3];
            __syncthreads();
            if (all) break;
#pragma unroll 1
            for (int p = 0; p < N; ++p)
#pragma unroll 1
                for (int q = p + 1; q < N; ++q) {
                    const double apq = A[p * N + q];
                    const bool rot = !gdone && apq != 0.0;
                    double c = 0.0, sn = 0.0;
                    if (rot) {
                        const double app = A[p * N + p], aqq = A[q * N + q];
                        const double theta = (aqq - app) / (2.0 * apq);
                        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                        c = 1.0 / sqrt(t * t + 1.0);
                        sn = t * c;
                    }
                    __syncthreads();
                    if (rot && l < N) {
                        const int k = l;
                        const double akp = A[k * N + p], akq = A[k * N + q];
                        A[k * N + p] = c * akp - sn * akq;
                        A[k * N + q] = sn * akp + c * akq;
                        const double vkp = V[k * N + p], vkq = V[k * N + q];
                        V[k * N + p] = c * vkp - sn * vkq;
                        V[k * N + q] = sn * vkp + c * vkq;
                    }
                    __syncthreads();
                    if (rot && l < N) {
                        const int k = l;
                        const double apk = A[p * N + k], aqk = A[q * N + k];
                        A[p * N + k] = c * apk - sn * aqk;
                        A[q * N + k] = sn * apk + c * aqk;
                    }
                    __syncthreads();
                }
        }
        if (active)
            for (int idx = l; idx < N * N; idx += kIpcLanes) {
                const int r = idx / N, c = idx % N;
                double s = 0.0;
                for (int k = 0; k < N; ++k) {
                    const double lk = A[k * N + k];
                    if (lk > 0.0) s += V[r * N + k] * lk * V[c * N + k];
                }
                sP[grp][idx] = s;
            }
        __syncthreads();
        Hs = sP[grp];
    }
    const int* own = sOwn[grp];
    const double* jth = sJ[grp];
    if (active) {
        if (l < N) {
            double s = 0.0;
            for (int r = 0; r < N; ++r) s += ipc_jel(r, l, own, jth) * sG[grp][r];
            pg[12 * size_t(i) + l] = s;
        }
        for (int idx = l; idx < N * N; idx += kIpcLanes) {
            const int r = idx / N, c = idx % N;
            double s = 0.0;
            for (int k = 0; k < N; ++k) s += Hs[r * N + k] * ipc_jel(k, c, own, jth);
            sV[grp][idx] = s;
        }
    }
    __syncthreads();
    if (active)
        for (int idx = l; idx < N * N; idx += kIpcLanes) {
            const int r = idx / N, c = idx % N;
            double s = 0.0;
            for (int k = 0; k < N; ++k) s += ipc_jel(k, r, own, jth) * sV[grp][k * N + c];
            pH[144 * size_t(i) + idx] = s;
        }
    __syncthreads();  // the next list entries reuse the LDS
    }
}

// Segments in three passes over chunks of kIpcChunk pairs: per chunk the number of
// segment starts (pair 0, or a key unlike the previous pair's) and of pairs within dhat;
// one work-group's exclusive scan of the chunk counts; per chunk the starts written in
// order. seg_start[s] = first pair of segment s (s < n_seg), seg_start[n_seg] = np;
// n_act = pairs within dhat.
constexpr int kIpcChunk = 1024;
constexpr int kIpcChunkThreads = 256;
constexpr int kIpcPerThread = kIpcChunk / kIpcChunkThreads;

__device__ __forceinline__ bool seg_first(const unsigned long long* key, int i) {
    return i == 0 || key[i] != key[i - 1];
}

__global__ __launch_bounds__(kIpcChunkThreads) void k_ipc_seg_count(int np, const unsigned long long* key,
                                                                    const int* act, int* chunk_cnt,
                                                                    int* chunk_act) {
    __shared__ int sc[kIpcChunkThreads], sa[kIpcChunkThreads];
    const int t = threadIdx.x, c0 = blockIdx.x * kIpcChunk;
    int c = 0, ca = 0;
    for (int k = 0; k < kIpcPerThread; ++k) {
        const int i = c0 + k * kIpcChunkThreads + t;
        if (i < np) {
            c += seg_first(key, i) ? 1 : 0;
            ca += act[i];
        }
    }
    sc[t] = c;
    sa[t] = ca;
    __syncthreads();
    for (int h = kIpcChunkThreads / 2; h > 0; h >>= 1) {
        if (t < h) {
            sc[t] += sc[t + h];
            sa[t] += sa[t + h];
        }
        __syncthreads();
    }
    if (t == 0) {
        chunk_cnt[blockIdx.x] = sc[0];
        chunk_act[blockIdx.x] = sa[0];
    }
}

// exclusive scan of the chunk counts in place (nchunk <= kIpcScanThreads * 4)
__global__ __launch_bounds__(kIpcScanThreads) void k_ipc_seg_scan(int np, int nchunk, int* chunk_cnt,
                                                                  int* chunk_act, int* seg_start, int* n_seg,
                                                                  int* n_act) {
    __shared__ int part[kIpcScanThreads], apart[kIpcScanThreads];
    const int t = threadIdx.x;
    constexpr int kPer = 4;
    // chunk_act is scanned as well: it becomes each chunk's offset in the active list
    int c = 0, ca = 0;
    for (int k = 0; k < kPer; ++k) {
        const int j = t * kPer + k;
        if (j < nchunk) {
            c += chunk_cnt[j];
            ca += chunk_act[j];
        }
    }
    part[t] = c;
    apart[t] = ca;
    __syncthreads();
    if (t == 0) {
        int run = 0, arun = 0;
        for (int k = 0; k < kIpcScanThreads; ++k) {
            const int v = part[k], va = apart[k];
            part[k] = run;
            apart[k] = arun;
            run += v;
            arun += va;
        }
        *n_seg = run;
        *n_act = arun;
        seg_start[run] = np;
    }
    __syncthreads();
    int run = part[t], arun = apart[t];
    for (int k = 0; k < kPer; ++k) {
        const int j = t * kPer + k;
        if (j < nchunk) {
            const int v = chunk_cnt[j], va = chunk_act[j];
            chunk_cnt[j] = run;
            chunk_act[j] = arun;
            run += v;
            arun += va;
        }
    }
}

__global__ __launch_bounds__(kIpcChunkThreads) void k_ipc_seg_write(int np, const unsigned long long* key,
                                                                    const int* act, const int* chunk_off,
                                                                    const int* chunk_aoff, int* seg_start,
                                                                    int* alist) {
    __shared__ int sc[kIpcChunkThreads], sa[kIpcChunkThreads];
    const int t = threadIdx.x, c0 = blockIdx.x * kIpcChunk + t * kIpcPerThread;
    int f[kIpcPerThread], fa[kIpcPerThread], c = 0, ca = 0;
    for (int k = 0; k < kIpcPerThread; ++k) {
        const int i = c0 + k;
        f[k] = (i < np && seg_first(key, i)) ? 1 : 0;
        fa[k] = (i < np && act[i] != 0) ? 1 : 0;
        c += f[k];
        ca += fa[k];
    }
    sc[t] = c;
    sa[t] = ca;
    __syncthreads();
    for (int h = 1; h < kIpcChunkThreads; h <<= 1) {  // inclusive scans (Hillis-Steele)
        const int v = t >= h ? sc[t - h] : 0, va = t >= h ? sa[t - h] : 0;
        __syncthreads();
        sc[t] += v;
        sa[t] += va;
        __syncthreads();
    }
    int pos = chunk_off[blockIdx.x] + sc[t] - c, apos = chunk_aoff[blockIdx.x] + sa[t] - ca;
    for (int k = 0; k < kIpcPerThread; ++k) {
        if (f[k]) seg_start[pos++] = c0 + k;
        if (fa[k]) alist[apos++] = c0 + k;
    }
}

// The pairs within dhat of segment s are a contiguous run of the in-order list alist:
// [first entry >= seg_start[s], first entry >= seg_start[s + 1]), found by binary search.
// Every output lane adds its element over that run in order; pairs out of dhat
// contribute nothing, so the sums are the same as over the whole segment in pair order.
__device__ __forceinline__ int lower_entry(const int* a, int n, int v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(kIpcReduceThreads) void k_ipc_reduce(const int* seg_start, const int* n_seg,
                                                                  const unsigned long long* key, const int* alist,
                                                                  const int* n_act, const double* pe,
                                                                  const double* pg, const double* pH, int value_only,
                                                                  int* seg_bodies, double* seg_E, double* seg_g,
                                                                  double* seg_H) {
    const int ns = *n_seg, na = *n_act;
    const int t = threadIdx.x;
    for (int s = blockIdx.x; s < ns; s += gridDim.x) {
        const int i0 = seg_start[s], i1 = seg_start[s + 1];
        const int k0 = lower_entry(alist, na, i0), k1 = lower_entry(alist, na, i1);
        double acc = 0.0;
        if (t == 0) {
            for (int k = k0; k < k1; ++k) acc += pe[alist[k]];
            seg_E[s] = acc;
            const unsigned long long kk = key[i0];
            seg_bodies[2 * s] = static_cast<int>(kk >> 32);
            seg_bodies[2 * s + 1] = static_cast<int>(kk & 0xffffffffull);
        } else if (value_only) {
        } else if (t <= 12) {
            for (int k = k0; k < k1; ++k) acc += pg[12 * size_t(alist[k]) + (t - 1)];
            seg_g[12 * size_t(s) + (t - 1)] = acc;
        } else if (t <= 156) {
            for (int k = k0; k < k1; ++k) acc += pH[144 * size_t(alist[k]) + (t - 13)];
            seg_H[144 * size_t(s) + (t - 13)] = acc;
        }
    }
}

// the segments' energies summed in segment order: staged through LDS in blocks (all
// lanes load), each block summed in order by one lane
__global__ __launch_bounds__(1024) void k_ipc_total(const int* n_seg, const double* seg_E, double* value) {
    __shared__ double se[1024];
    const int ns = *n_seg;
    double v = 0.0;
    for (int s0 = 0; s0 < ns; s0 += 1024) {
        const int m = ns - s0 < 1024 ? ns - s0 : 1024;
        if (static_cast<int>(threadIdx.x) < m) se[threadIdx.x] = seg_E[s0 + threadIdx.x];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int k = 0; k < m; ++k) v += se[k];
        __syncthreads();
    }
    if (threadIdx.x == 0) *value = v;
}

// ipc mode, OptimizationProblem::maxStep (NewtonSolver.h:42, NewtonSolver.cpp:53-54),
// one lane per pair: conservative advancement (ipc::pair_advance) along the line-search
// path x + t dx (positions p + t dp, orientations apply_theta(q, theta + t dtheta)),
// keeping sep of the pair's current distance. mu = sum over the two bodies of |dp| +
// |dtheta| r_max (r_max: the largest body-frame radius of the pair's vertices on that
// body), margin = 2e-10 (r_max A + r_max B) for the rotation skipped below |theta| =
// 1e-10. toi[i] = the pair's advance (+inf: no impact in the step), or where it stopped
// once it could no longer be the least; st[i] = -3 for a bad pair.
constexpr int kIpcAdvanceIters = 1000;

__global__ __launch_bounds__(64) void k_ipc_advance(Dev d, int b0, int nb, int np, const int* pairs, const double* x,
                                                    const double* dx, double sep, double* toi, int* st,
                                                    unsigned long long* best_bits) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const int* pr = pairs + 5 * i;
    const int kind = pr[0], a = pr[1], b = pr[2], ia = pr[3], ib = pr[4];
    bool ok = (kind == 0 || kind == 1) && a >= 0 && b >= 0 && a < nb && b < nb && a != b;
    const int ga = ok ? b0 + a : 0, gb = ok ? b0 + b : 0;
    ok = ok && d.nv[ga] > 0 && d.nv[gb] > 0 && ia >= 0 && ib >= 0;
    ok = ok && (kind == 0 ? (ia < d.nv[ga] && ib < d.nf[gb]) : (ia < d.ne[ga] && ib < d.ne[gb]));
    if (!ok) {
        toi[i] = 1.0 / 0.0;
        st[i] = -3;
        return;
    }
    V3 r[4];
    if (kind == 0) {
        r[0] = ld3(d.verts, d.v_off[ga] + ia);
        const int* f = d.faces + 3 * (d.f_off[gb] + ib);
#pragma unroll
        for (int k = 0; k < 3; ++k) r[1 + k] = ld3(d.verts, d.v_off[gb] + f[k]);
    } else {
        const int* ea = d.edges + 2 * (d.e_off[ga] + ia);
        const int* eb = d.edges + 2 * (d.e_off[gb] + ib);
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            r[k] = ld3(d.verts, d.v_off[ga] + ea[k]);
            r[2 + k] = ld3(d.verts, d.v_off[gb] + eb[k]);
        }
    }
    // every per-vertex quantity in registers (unrolled, fixed indices): vertex v is on
    // body a for v < na_ (one vertex of a point-triangle pair, two of an edge-edge pair)
    const int na_ = kind == 0 ? 1 : 2;
    double rmax_a = 0.0, rmax_b = 0.0;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const double rl = sqrt((r[v].x * r[v].x + r[v].y * r[v].y) + r[v].z * r[v].z);
        if (v < na_) rmax_a = dmax(rmax_a, rl);
        else rmax_b = dmax(rmax_b, rl);
    }
    const double rmax[2] = {rmax_a, rmax_b};
    double mu = 0.0;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const int o = 6 * (side == 0 ? a : b);
        const double lp = sqrt((dx[o] * dx[o] + dx[o + 1] * dx[o + 1]) + dx[o + 2] * dx[o + 2]);
        const double lt = sqrt((dx[o + 3] * dx[o + 3] + dx[o + 4] * dx[o + 4]) + dx[o + 5] * dx[o + 5]);
        mu = mu + (lp + lt * rmax[side]);
    }
    const double margin = 2e-10 * (rmax[0] + rmax[1]);
    const Q qa = ld4(d.q, ga), qb = ld4(d.q, gb);
    const int oa = 6 * a, ob = 6 * b;
    auto world = [&](double t, double* xw) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const bool on_a = v < na_;
            const int o = on_a ? oa : ob;
            const V3 p = v3(x[o] + t * dx[o], x[o + 1] + t * dx[o + 1], x[o + 2] + t * dx[o + 2]);
            const V3 th = v3(x[o + 3] + t * dx[o + 3], x[o + 4] + t * dx[o + 4], x[o + 5] + t * dx[o + 5]);
            const V3 w = p + rotate(apply_theta(on_a ? qa : qb, th), r[v]);
            xw[3 * v] = w.x;
            xw[3 * v + 1] = w.y;
            xw[3 * v + 2] = w.z;
        }
    };
    // a pair whose advance has reached the least finished advance so far cannot be the
    // least: it stops there (its true advance is no smaller), which leaves the least
    // over all pairs unchanged
    const volatile unsigned long long* bb = best_bits;
    auto done = [&](double t) { return !(t < __longlong_as_double(static_cast<long long>(*bb))); };
    const double t = ipc::pair_advance(kind, world, mu, sep, margin, kIpcAdvanceIters, done);
    toi[i] = t;
    st[i] = 0;
    if (t < 1.0) atomicMin(best_bits, static_cast<unsigned long long>(__double_as_longlong(t)));
}

// least advance and first bad pair over every pair: a grid-wide min through 64-bit
// atomics on the advances' bit patterns (non-negative doubles, +inf included, order as
// their bits do), exact in any order; *out_bits starts at the bits of +inf, *out_bad at np
__global__ __launch_bounds__(256) void k_ipc_advance_min(int np, const double* toi, const int* st,
                                                         unsigned long long* out_bits, int* out_bad) {
    __shared__ unsigned long long sm[256];
    __shared__ int sb[256];
    unsigned long long m = 0x7ff0000000000000ull;
    int fb = np;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < np; i += gridDim.x * 256) {
        const unsigned long long b = __double_as_longlong(toi[i]);
        m = b < m ? b : m;
        if (st[i] == -3 && i < fb) fb = i;
    }
    sm[threadIdx.x] = m;
    sb[threadIdx.x] = fb;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            sm[threadIdx.x] = sm[threadIdx.x + h] < sm[threadIdx.x] ? sm[threadIdx.x + h] : sm[threadIdx.x];
            sb[threadIdx.x] = min(sb[threadIdx.x], sb[threadIdx.x + h]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicMin(out_bits, sm[0]);
        atomicMin(out_bad, sb[0]);
    }
}

void launch_ipc_advance(const Dev& d, int b0, int nb, int np, const int* pairs, const double* x, const double* dx,
                        double sep, double* toi, int* st, double* out_t, int* out_bad, hipStream_t stream) {
    if (np <= 0) return;
    timing_begin("ipc_advance", stream);
    hipLaunchKernelGGL(k_ipc_advance, dim3((np + 63) / 64), dim3(64), 0, stream, d, b0, nb, np, pairs, x, dx, sep, toi,
                       st, reinterpret_cast<unsigned long long*>(out_t));
    const int grid = (np + 255) / 256 < 1024 ? (np + 255) / 256 : 1024;
    hipLaunchKernelGGL(k_ipc_advance_min, dim3(grid), dim3(256), 0, stream, np, toi, st,
                       reinterpret_cast<unsigned long long*>(out_t), out_bad);
    timing_end("ipc_advance", stream);
}

void launch_ipc_eval(const Dev& d, const IpcEvalBufs& B, int b0, int nb, int np, double dhat, double kappa, int psd,
                     hipStream_t st, int value_only) {
    if (np <= 0) return;
    const int nchunk = (np + kIpcChunk - 1) / kIpcChunk;
    timing_begin("ipc_filter", st);
    hipLaunchKernelGGL(k_ipc_filter, dim3((np + 255) / 256), dim3(256), 0, st, d, b0, nb, np, B.pairs, B.x, dhat,
                       kappa, value_only, B.key, B.act, B.bad, B.pe);
    hipLaunchKernelGGL(k_ipc_seg_count, dim3(nchunk), dim3(kIpcChunkThreads), 0, st, np, B.key, B.act, B.chunk_cnt,
                       B.chunk_act);
    hipLaunchKernelGGL(k_ipc_seg_scan, dim3(1), dim3(kIpcScanThreads), 0, st, np, nchunk, B.chunk_cnt, B.chunk_act,
                       B.seg_start, B.n_seg, B.n_act);
    hipLaunchKernelGGL(k_ipc_seg_write, dim3(nchunk), dim3(kIpcChunkThreads), 0, st, np, B.key, B.act, B.chunk_cnt,
                       B.chunk_act, B.seg_start, B.alist);
    timing_end("ipc_filter", st);
    if (!value_only) {
        timing_begin("ipc_pairs", st);
        const int gp = (np + kIpcPairsPerWg - 1) / kIpcPairsPerWg;
        hipLaunchKernelGGL(k_ipc_pairs, dim3(gp < 2048 ? gp : 2048), dim3(64), 0, st, d, b0, B.pairs, B.x, dhat, kappa,
                           psd, B.alist, B.n_act, B.pe, B.pg, B.pH);
        timing_end("ipc_pairs", st);
    }
    timing_begin("ipc_reduce", st);
    const int grid = np < 4096 ? np : 4096;
    hipLaunchKernelGGL(k_ipc_reduce, dim3(grid), dim3(kIpcReduceThreads), 0, st, B.seg_start, B.n_seg, B.key, B.alist,
                       B.n_act, B.pe, B.pg, B.pH, value_only, B.seg_bodies, B.seg_E, B.seg_g, B.seg_H);
    hipLaunchKernelGGL(k_ipc_total, dim3(1), dim3(1024), 0, st, B.n_seg, B.seg_E, B.value);
    timing_end("ipc_reduce", st);
}

// ------------------------------------------------------------------ ipc-mode step
// Per close pair of the scene: the number of pair records of its top-K lists (faces of A
// x vertices of B, faces of B x vertices of A, edges of A x edges of B), 0 when both
// bodies are static.
__global__ void k_ipc_emit_count(int ncl, const int* cpairs, const int* lcnt, const int* stat, int b0, int* cnt) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= ncl) return;
    const int A = cpairs[2 * p] - b0, B = cpairs[2 * p + 1] - b0;
    const int* k = lcnt + 6 * p;
    cnt[p] = (stat[A] && stat[B]) ? 0 : k[4] * k[1] + k[5] * k[0] + k[2] * k[3];
}

// One work-group per close pair writes its records from off[p], in the order above.
__global__ void k_ipc_emit_write(int ncl, const int* cpairs, const int* lists, const int* lcnt, const int* off,
                                 const int* cnt, int b0, int* out) {
    const int p = blockIdx.x;
    if (p >= ncl) return;
    const int A = cpairs[2 * p] - b0, B = cpairs[2 * p + 1] - b0;
    const int* l = lists + size_t(p) * 6 * kTopK;
    const int* k = lcnt + 6 * p;
    const int n1 = k[4] * k[1], n2 = k[5] * k[0];
    for (int t = threadIdx.x; t < cnt[p]; t += blockDim.x) {
        int* o = out + 5 * size_t(off[p] + t);
        if (t < n1) {
            const int i = t / k[1], j = t % k[1];
            o[0] = 0; o[1] = B; o[2] = A; o[3] = l[kTopK + j]; o[4] = l[4 * kTopK + i];
        } else if (t < n1 + n2) {
            const int u = t - n1, i = u / k[0], j = u % k[0];
            o[0] = 0; o[1] = A; o[2] = B; o[3] = l[j]; o[4] = l[5 * kTopK + i];
        } else {
            const int u = t - n1 - n2, i = u / k[3], j = u % k[3];
            o[0] = 1; o[1] = A; o[2] = B; o[3] = l[2 * kTopK + i]; o[4] = l[3 * kTopK + j];
        }
    }
}

// x = (p, 0), x^ = (p + h v, h w) per body of the scene
__global__ void k_ipc_start(Dev d, int b0, int nb, double h, double* x, double* xh) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const int g = b0 + b;
    for (int k = 0; k < 3; ++k) {
        x[6 * b + k] = d.p[3 * g + k];
        x[6 * b + 3 + k] = 0.0;
        xh[6 * b + k] = d.p[3 * g + k] + h * d.v[3 * g + k];
        xh[6 * b + 3 + k] = h * d.w[3 * g + k];
    }
}

// Per body of the scene: its inertia energy and gravity term (ipc::body_terms); unless
// energy_only, its gradient and diagonal block: inertia, h^2 x the gravity force, then
// h^2 x its parts of the barrier segments in segment order. Static bodies: identity
// block, zero gradient, no energy.
__global__ void k_ipc_body(Dev d, int b0, int nb, const int* stat, const double* x, const double* xh, double h2,
                           int energy_only, const int* n_seg, const int* seg_bodies, const double* seg_g,
                           const double* seg_H, double* g, double* diag, double* ebody, double* gterm) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const int gb = b0 + b;
    double gl[6], D[36];
    if (stat[b]) {
        ebody[b] = 0.0;
        gterm[b] = 0.0;
        if (energy_only) return;
        for (int k = 0; k < 36; ++k) diag[36 * b + k] = (k % 7 == 0) ? 1.0 : 0.0;
        for (int k = 0; k < 6; ++k) g[6 * b + k] = 0.0;
        return;
    }
    double e, sgd = 0.0;
    const double m = d.mass[gb];
    ipc::body_terms(m, d.inertia + 9 * gb, x + 6 * b, xh + 6 * b, e, gl, D, &sgd);
    ebody[b] = e;
    gterm[b] = sgd;
    if (energy_only) return;
    for (int k = 0; k < 3; ++k) gl[k] += h2 * (-m * ipc::gravity(k));
    const int ns = *n_seg;
    for (int sg = 0; sg < ns; ++sg) {
        const int lo = seg_bodies[2 * sg], hi = seg_bodies[2 * sg + 1];
        const int side = lo == b ? 0 : (hi == b ? 1 : -1);
        if (side < 0) continue;
        for (int r = 0; r < 6; ++r) {
            gl[r] += h2 * seg_g[12 * size_t(sg) + 6 * side + r];
            for (int c = 0; c < 6; ++c)
                D[6 * r + c] += h2 * seg_H[144 * size_t(sg) + 12 * (6 * side + r) + 6 * side + c];
        }
    }
    for (int k = 0; k < 6; ++k) g[6 * b + k] = gl[k];
    for (int k = 0; k < 36; ++k) diag[36 * b + k] = D[k];
}

// one off-diagonal block per segment: h^2 x its (lower, higher) block, zero when either
// body is static; and rhs = -g
__global__ void k_ipc_offblocks(int nd, const int* stat, double h2, const int* n_seg, const int* seg_bodies,
                                const double* seg_H, const double* g, int* offij, double* offb, double* rhs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nd) rhs[i] = -g[i];
    const int ns = *n_seg;
    if (i >= ns) return;
    const int lo = seg_bodies[2 * i], hi = seg_bodies[2 * i + 1];
    offij[2 * i] = lo;
    offij[2 * i + 1] = hi;
    const bool both = !stat[lo] && !stat[hi];
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c)
            offb[36 * size_t(i) + 6 * r + c] = both ? h2 * seg_H[144 * size_t(i) + 12 * r + 6 + c] : 0.0;
}

// out[0] = sum of the moving bodies' inertia energies, out[1] = of their gravity terms
// (body order); with dx: out[2] = g . dx (in order), out[3] = |dx|_inf
__global__ void k_ipc_sums(int nb, const int* stat, const double* ebody, const double* gterm, const double* g,
                           const double* dx, double* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double si = 0.0, sg = 0.0;
    for (int b = 0; b < nb; ++b) {
        if (stat[b]) continue;
        si += ebody[b];
        sg += gterm[b];
    }
    out[0] = si;
    out[1] = sg;
    if (!dx) return;
    double slope = 0.0, dmax = 0.0;
    for (int i = 0; i < 6 * nb; ++i) {
        slope += g[i] * dx[i];
        dmax = dmax < fabs(dx[i]) ? fabs(dx[i]) : dmax;
    }
    out[2] = slope;
    out[3] = dmax;
}

__global__ void k_ipc_axpy(int nd, const double* x, double alpha, const double* dx, double* xt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nd) xt[i] = x[i] + alpha * dx[i];
}

// write-back of the moving bodies (IPCSolver::step): p = x_p, v = (x_p - p) / h,
// w = x_theta / h, q = apply_theta(q, x_theta)
__global__ void k_ipc_writeback(Dev d, int b0, int nb, const int* stat, const double* x, double h) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb || stat[b]) return;
    const int g = b0 + b;
    const V3 pn = v3(x[6 * b], x[6 * b + 1], x[6 * b + 2]), th = v3(x[6 * b + 3], x[6 * b + 4], x[6 * b + 5]);
    const V3 p0 = ld3(d.p, g);
    const V3 v = (pn - p0) / h, w = th / h;
    d.v[3 * g] = v.x; d.v[3 * g + 1] = v.y; d.v[3 * g + 2] = v.z;
    d.w[3 * g] = w.x; d.w[3 * g + 1] = w.y; d.w[3 * g + 2] = w.z;
    d.p[3 * g] = pn.x; d.p[3 * g + 1] = pn.y; d.p[3 * g + 2] = pn.z;
    const Q q = apply_theta(ld4(d.q, g), th);
    d.q[4 * g] = q.w; d.q[4 * g + 1] = q.x; d.q[4 * g + 2] = q.y; d.q[4 * g + 3] = q.z;
}

void launch_ipc_emit(const Work& w, int ncl, const int* stat, int b0, int* cnt, const int* off, int* out, int phase,
                     hipStream_t st) {
    if (ncl <= 0) return;
    if (phase == 0)
        hipLaunchKernelGGL(k_ipc_emit_count, dim3((ncl + 255) / 256), dim3(256), 0, st, ncl, w.pairs, w.lcnt, stat, b0,
                           cnt);
    else
        hipLaunchKernelGGL(k_ipc_emit_write, dim3(ncl), dim3(256), 0, st, ncl, w.pairs, w.lists, w.lcnt, off, cnt, b0,
                           out);
}
void launch_ipc_start(const Dev& d, int b0, int nb, double h, double* x, double* xh, hipStream_t st) {
    hipLaunchKernelGGL(k_ipc_start, dim3((nb + 255) / 256), dim3(256), 0, st, d, b0, nb, h, x, xh);
}
void launch_ipc_body(const Dev& d, int b0, int nb, const int* stat, const double* x, const double* xh, double h2,
                     int energy_only, const IpcEvalBufs& B, double* g, double* diag, double* ebody, double* gterm,
                     hipStream_t st) {
    hipLaunchKernelGGL(k_ipc_body, dim3((nb + 63) / 64), dim3(64), 0, st, d, b0, nb, stat, x, xh, h2, energy_only,
                       B.n_seg, B.seg_bodies, B.seg_g, B.seg_H, g, diag, ebody, gterm);
}
void launch_ipc_offblocks(int nd, int nseg_max, const int* stat, double h2, const IpcEvalBufs& B, const double* g,
                          int* offij, double* offb, double* rhs, hipStream_t st) {
    const int n = nd > nseg_max ? nd : nseg_max;
    hipLaunchKernelGGL(k_ipc_offblocks, dim3((n + 255) / 256), dim3(256), 0, st, nd, stat, h2, B.n_seg, B.seg_bodies,
                       B.seg_H, g, offij, offb, rhs);
}
void launch_ipc_sums(int nb, const int* stat, const double* ebody, const double* gterm, const double* g,
                     const double* dx, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_ipc_sums, dim3(1), dim3(64), 0, st, nb, stat, ebody, gterm, g, dx, out);
}
void launch_ipc_axpy(int nd, const double* x, double alpha, const double* dx, double* xt, hipStream_t st) {
    hipLaunchKernelGGL(k_ipc_axpy, dim3((nd + 255) / 256), dim3(256), 0, st, nd, x, alpha, dx, xt);
}
void launch_ipc_writeback(const Dev& d, int b0, int nb, const int* stat, const double* x, double h, hipStream_t st) {
    hipLaunchKernelGGL(k_ipc_writeback, dim3((nb + 255) / 256), dim3(256), 0, st, d, b0, nb, stat, x, h);
}

}  // namespace nd
