// Contact-stage geometry kernels for gfx950: world AABBs, per-scene LBVH broad phase,
// tie-exact top-K candidate lists, narrow phase with order-preserving compaction, the
// scene-minimum-distance flag and interval CCD.
//
// All kernels are FP64 VALU/gather work (no MFMA). Work is flattened over
// scene x pair so one launch serves every scene of a batch; per-pair kernels run a
// persistent grid so the device-side pair count needs no host round trip.
#include "ndipc_internal.h"
#include "topk_select.h"

#include <algorithm>

namespace nd {

constexpr int kBlock = 256;

// ------------------------------------------------------------------- world boxes
// ContactAssembler::build -> transform_local_aabb (ContactAssembler.cpp:22-43,145-149).
__global__ void k_world_boxes(Dev d, Work w) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= d.N) return;
    if (d.nv[b] == 0) return;
    const Box3 lb = ldbox(d.lbox, b);
    const Box3 wb = world_box(lb.mn, lb.mx, ld3(d.p, b), ld4(d.q, b));
    double* o = w.wbox + 6 * b;
    o[0] = wb.mn.x; o[1] = wb.mn.y; o[2] = wb.mn.z;
    o[3] = wb.mx.x; o[4] = wb.mx.y; o[5] = wb.mx.z;
}

void launch_world_boxes(const Dev& d, const Work& w, hipStream_t st) {
    if (d.N == 0) return;
    hipLaunchKernelGGL(k_world_boxes, dim3((d.N + kBlock - 1) / kBlock), dim3(kBlock), 0, st, d, w);
}

// ------------------------------------------------------------------------- LBVH
// Karras 2012 ("Maximizing parallelism in the construction of BVHs, octrees and k-d
// trees") over 30-bit Morton codes of the body AABB centres, one work-group per scene.
// Keys carry the leaf index in the low 32 bits, so they are unique. Internal-node boxes
// are the union of the node's contiguous leaf range.
__device__ __forceinline__ unsigned expand_bits(unsigned v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__device__ __forceinline__ int karras_delta(const unsigned long long* keys, int m, int i, int j) {
    if (j < 0 || j >= m) return -1;
    return __clzll(keys[i] ^ keys[j]);
}

// Every query of the tree returns the set of leaves whose boxes overlap the query box,
// and every consumer orders its pairs itself (j descending per i), so results do not
// depend on the tree's shape: the Morton sort and radix tree are rebuilt every
// kBvhRebuild-th call of a scene, and the calls in between only refit the node boxes to
// the current body boxes over the kept topology.
constexpr int kBvhRebuild = 16;

__device__ void lbvh_node_boxes(const Work& w, int s, int b0, int m);

__global__ void __launch_bounds__(kBlock) k_lbvh_build(Dev d, Work w) {
    __shared__ unsigned long long keys[kMaxSceneBodies];
    __shared__ int leaf_tmp[kMaxSceneBodies];
    __shared__ double red[kBlock];
    __shared__ int redi[kBlock];
    const int s = blockIdx.x;
    const int t = threadIdx.x;
    const int b0 = d.body_off[s], b1 = d.body_off[s + 1];
    const int nb = b1 - b0;
    const int age = w.bvh_age[s];
    __syncthreads();
    if (t == 0) w.bvh_age[s] = age + 1 == kBvhRebuild ? 0 : age + 1;
    if (age != 0) {  // refit
        const int m = w.nleaves[s];
        if (m > 1) lbvh_node_boxes(w, s, b0, m);
        return;
    }
    // compact the mesh bodies (scene order)
    int m = 0;
    for (int base = 0; base < nb; base += kBlock) {
        const int i = base + t;
        const int has = (i < nb && d.nv[b0 + i] > 0) ? 1 : 0;
        int ex;
        const int tot = wg_exclusive_scan(has, redi, &ex);
        if (has) leaf_tmp[m + ex] = b0 + i;
        m += tot;
        __syncthreads();
    }
    if (t == 0) w.nleaves[s] = m;
    if (m == 0) return;
    // centroid bounds
    double lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        double vmn = 1.7976931348623157e308, vmx = -1.7976931348623157e308;
        for (int k = t; k < m; k += kBlock) {
            const double* bx = w.wbox + 6 * leaf_tmp[k];
            const double c = (bx[a] + bx[3 + a]) * 0.5;
            vmn = dmin(vmn, c);
            vmx = dmax(vmx, c);
        }
        lo[a] = wg_reduce(vmn, red, OpMin());
        hi[a] = wg_reduce(vmx, red, OpMax());
    }
    int P = 1;
    while (P < m) P <<= 1;
    for (int k = t; k < P; k += kBlock) {
        if (k < m) {
            const double* bx = w.wbox + 6 * leaf_tmp[k];
            unsigned code = 0;
            unsigned c3[3];
            for (int a = 0; a < 3; ++a) {
                const double c = (bx[a] + bx[3 + a]) * 0.5;
                const double ext = hi[a] - lo[a];
                double u = ext > 0.0 ? (c - lo[a]) / ext : 0.0;
                u = dmin(dmax(u, 0.0), 1.0);
                c3[a] = (unsigned)dmin(u * 1024.0, 1023.0);
            }
            code = (expand_bits(c3[0]) << 2) | (expand_bits(c3[1]) << 1) | expand_bits(c3[2]);
            keys[k] = ((unsigned long long)code << 32) | (unsigned)k;
        } else {
            keys[k] = ~0ull;
        }
    }
    __syncthreads();
    // bitonic sort of the keys in LDS
    for (int size = 2; size <= P; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int k = t; k < P / 2; k += kBlock) {
                const int i = 2 * k - (k & (stride - 1));
                const int j = i + stride;
                const bool up = ((i & size) == 0);
                const unsigned long long a = keys[i], b = keys[j];
                if ((a > b) == up) { keys[i] = b; keys[j] = a; }
            }
            __syncthreads();
        }
    }
    for (int k = t; k < m; k += kBlock) w.leaf_body[b0 + k] = leaf_tmp[keys[k] & 0xffffffffu];
    __syncthreads();
    // internal nodes 0 .. m-2
    for (int i = t; i < m - 1; i += kBlock) {
        const int dl = karras_delta(keys, m, i, i + 1) - karras_delta(keys, m, i, i - 1);
        const int dir = dl > 0 ? 1 : -1;
        const int dmin_ = karras_delta(keys, m, i, i - dir);
        int lmax = 2;
        while (karras_delta(keys, m, i, i + lmax * dir) > dmin_) lmax <<= 1;
        int l = 0;
        for (int tt = lmax >> 1; tt >= 1; tt >>= 1)
            if (karras_delta(keys, m, i, i + (l + tt) * dir) > dmin_) l += tt;
        const int j = i + l * dir;
        const int dnode = karras_delta(keys, m, i, j);
        int sp = 0, tt = l;
        do {
            tt = (tt + 1) >> 1;
            if (karras_delta(keys, m, i, i + (sp + tt) * dir) > dnode) sp += tt;
        } while (tt > 1);
        const int gamma = i + sp * dir + (dir < 0 ? -1 : 0);
        const int first = i < j ? i : j, last = i < j ? j : i;
        w.node_child[2 * (b0 + i)] = (first == gamma) ? ~gamma : gamma;
        w.node_child[2 * (b0 + i) + 1] = (last == gamma + 1) ? ~(gamma + 1) : gamma + 1;
        w.node_range[2 * (b0 + i)] = first;
        w.node_range[2 * (b0 + i) + 1] = last;
    }
    __syncthreads();
    lbvh_node_boxes(w, s, b0, m);
}

// Node boxes: the union of the node's leaf range [first, last], from a sparse table of
// power-of-two range unions in the scene's scratch: T_0[k] = box of sorted leaf k,
// T_l[k] = T_{l-1}[k] u T_{l-1}[k + 2^(l-1)], built level by level; a range is then
// covered by steps of 2^l from `first` plus one entry ending at `last` (l the largest
// level stored not above log2 of its length). Min and max are exact, so the box does
// not depend on how the range is covered. Called by the whole work-group, m > 1.
__device__ void lbvh_node_boxes(const Work& w, int s, int b0, int m) {
    const int t = threadIdx.x;
    double* tab = w.sort_key + (long long)s * kSceneScratch;
    int L = 0;
    while ((2 << L) <= m && 6 * m * (L + 2) <= kSceneScratch) ++L;  // levels 0..L
    for (int k = t; k < m; k += kBlock) {
        const double* bx = w.wbox + 6 * w.leaf_body[b0 + k];
        for (int q = 0; q < 6; ++q) tab[6 * k + q] = bx[q];
    }
    __syncthreads();
    for (int l = 1; l <= L; ++l) {
        const double* pv = tab + 6 * m * (l - 1);
        double* cur = tab + 6 * m * l;
        const int h = 1 << (l - 1);
        for (int k = t; k + 2 * h <= m; k += kBlock) {
            for (int q = 0; q < 3; ++q) cur[6 * k + q] = dmin(pv[6 * k + q], pv[6 * (k + h) + q]);
            for (int q = 3; q < 6; ++q) cur[6 * k + q] = dmax(pv[6 * k + q], pv[6 * (k + h) + q]);
        }
        __syncthreads();
    }
    for (int i = t; i < m - 1; i += kBlock) {
        const int first = w.node_range[2 * (b0 + i)], last = w.node_range[2 * (b0 + i) + 1];
        const int len = last - first + 1;
        int l = 0;
        while (l < L && (2 << l) <= len) ++l;
        const int S = 1 << l;
        const double* tl = tab + 6 * m * l;
        double bx[6];
        for (int q = 0; q < 6; ++q) bx[q] = tl[6 * (last - S + 1) + q];
        for (int p = first; p + S - 1 < last; p += S) {
            for (int q = 0; q < 3; ++q) bx[q] = dmin(bx[q], tl[6 * p + q]);
            for (int q = 3; q < 6; ++q) bx[q] = dmax(bx[q], tl[6 * p + q]);
        }
        double* o = w.node_box + 6 * (b0 + i);
        for (int q = 0; q < 6; ++q) o[q] = bx[q];
    }
}

void launch_lbvh_build(const Dev& d, const Work& w, hipStream_t st) {
    timing_begin("lbvh_build", st);
    hipLaunchKernelGGL(k_lbvh_build, dim3(d.S), dim3(kBlock), 0, st, d, w);
    timing_end("lbvh_build", st);
}

// Traversal: leaf test = AABB::intersects of both boxes inflated by r
// (ContactAssembler.cpp:165, AABB.h:149-153,193-196), evaluated with the same FP
// expressions. Internal nodes use the identical expression on the union box; rounding
// is monotonic, so a rejected node cannot contain an accepted leaf.
template <typename F>
__device__ void lbvh_query(const Dev& d, const Work& w, int s, int i, double r, F&& on_hit) {
    const int b0 = d.body_off[s];
    const int m = w.nleaves[s];
    const Box3 qa = box_inflated(ldbox(w.wbox, i), r);
    if (m == 1) return;
    int stack[72];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const int node = stack[--sp];
        const int c0 = w.node_child[2 * (b0 + node)], c1 = w.node_child[2 * (b0 + node) + 1];
        const int cs[2] = {c0, c1};
        for (int k = 0; k < 2; ++k) {
            const int c = cs[k];
            if (c < 0) {
                const int j = w.leaf_body[b0 + (~c)];
                if (j > i && box_overlap(qa, box_inflated(ldbox(w.wbox, j), r))) on_hit(j);
            } else {
                if (box_overlap(qa, box_inflated(ldbox(w.node_box, b0 + c), r)) && sp < 72) stack[sp++] = c;
            }
        }
    }
}

// Fused broad phase of IPCSolver::step: the three detectBodyPairs queries of a step
// (candidates at the pre-initialisation dhat, ContactAssembler.cpp:220; stiffness
// check at 2 dhat, IPCSolver.cpp:825; CCD gate at max(1e-6, dhat) + v_max dt,
// IPCSolver.cpp:365) share one traversal per body with the largest inflation; each hit
// is then classified with the exact per-query expression. One thread per body over all
// scenes; pass 0 counts, pass 1 emits at scanned offsets.
//
// Pair order is the reference's: BVH::query walks its tree depth first with a stack,
// pushing the left child before the right one, and its builder never reorders the
// primitives (the SAH split index always lies strictly inside the range), so leaves are
// visited in descending primitive order. detectBodyPairs therefore lists (i, j) by i
// ascending, then j descending (BroadPhase.cpp, BVH.cpp query/build). The contact list,
// and with it every per-body sum of the Newton objective, inherits this order.
__device__ __forceinline__ void bp_sort_insert(int* outp, int o, int k, int i, int j) {
    int pos = k;  // entries [0, k) hold j descending; j goes to its place
    while (pos > 0 && outp[2 * (o + pos - 1) + 1] < j) {
        outp[2 * (o + pos)] = outp[2 * (o + pos - 1)];
        outp[2 * (o + pos) + 1] = outp[2 * (o + pos - 1) + 1];
        --pos;
    }
    outp[2 * (o + pos)] = i;
    outp[2 * (o + pos) + 1] = j;
}

// The scene's tree staged in LDS for the traversals of k_bp_step: internal nodes
// 0..m-2 (children, union boxes) and the m leaves in sorted order (body index, body id,
// world box), 112 bytes per leaf. Reads give the same values as the tree in HBM.
struct BvhLds {
    int* child;     // [2 (m - 1)]
    double* nbox;   // [6 (m - 1)]
    int* lbody;     // [m]
    int* lbid;      // [m] scene-local id of the leaf's body
    double* lbox;   // [6 m]
    unsigned short* stk;  // traversal stacks, kBvhStackBytes
    unsigned short* hit;  // per-lane hit buffers (nhits entries each)
};
constexpr int kBvhLdsPerLeaf = 2 * 4 + 6 * 8 + 4 + 4 + 6 * 8;
constexpr int kBvhStackBytes = 72 * kBlock * 2;  // 72 entries of 2 bytes per lane
constexpr int kBvhMinHits = 32;                  // buffered hits per lane, at least

__device__ __forceinline__ Box3 box_at(const double* b, int i) {
    Box3 r;
    r.mn = v3(b[6 * i], b[6 * i + 1], b[6 * i + 2]);
    r.mx = v3(b[6 * i + 3], b[6 * i + 4], b[6 * i + 5]);
    return r;
}

// lbvh_query over the LDS copy: the same traversal and tests. The node stack of each
// lane lives in LDS as well (16-bit node indices, entry e of lane t at stk[e * kBlock
// + t]): a stack in registers with a dynamic index would be placed in scratch memory.
// Accepted leaves are buffered per lane and handed to on_leaf(leaf) when the buffer is
// full and after the traversal, so the lanes of a wavefront process their hits in the
// same few iterations instead of one divergent call per hit. Callers must not depend
// on the order of the calls (pass 0 counts and takes minima, pass 1 inserts sorted).
// With `sorted`, each batch of buffered leaves is handed over by body index descending
// (an insertion sort of the lane's buffer), the order of the emitted pair lists.
constexpr int kBvhStack = 72;
template <typename F>
__device__ void lbvh_query_lds(const BvhLds& T, int m, int i, const Box3& bi, double r, bool sorted, int nhits,
                               F&& on_leaf) {
    const Box3 qa = box_inflated(bi, r);
    if (m == 1) return;
    unsigned short* st = T.stk + threadIdx.x;
    unsigned short* hb = T.hit + threadIdx.x;
    int nh = 0;
    int sp = 1;
    st[0] = 0;
    auto flush = [&]() {
        if (sorted) {
            for (int h = 1; h < nh; ++h) {
                const unsigned short v = hb[kBlock * h];
                const int key = T.lbody[v];
                int p = h;
                while (p > 0 && T.lbody[hb[kBlock * (p - 1)]] < key) {
                    hb[kBlock * p] = hb[kBlock * (p - 1)];
                    --p;
                }
                hb[kBlock * p] = v;
            }
        }
        for (int h = 0; h < nh; ++h) on_leaf(hb[kBlock * h]);
        nh = 0;
    };
    while (sp > 0) {
        const int node = st[kBlock * --sp];
        const int cs[2] = {T.child[2 * node], T.child[2 * node + 1]};
        for (int k = 0; k < 2; ++k) {
            const int c = cs[k];
            if (c < 0) {
                if (T.lbody[~c] > i && box_overlap(qa, box_inflated(box_at(T.lbox, ~c), r))) {
                    hb[kBlock * nh++] = (unsigned short)(~c);
                    if (nh == nhits) flush();
                }
            } else {
                if (box_overlap(qa, box_inflated(box_at(T.nbox, c), r)) && sp < kBvhStack)
                    st[kBlock * sp++] = (unsigned short)c;
            }
        }
    }
    flush();
}

// Grid (scene, chunk of kBlock bodies). With `lds` the work-group first copies its
// scene's tree into dynamic LDS (kBvhLdsPerLeaf bytes per leaf, sized for the largest
// scene at launch), followed by the lanes' stacks and hit buffers of nhits entries;
// otherwise the traversal reads the tree in HBM. The counting pass (pass 0) keeps each
// body's candidate and prox hits in its w.bp_cache slot (candidates from the front, prox
// from the back) when they fit in kBpCache entries; pass 1 then emits them (each list
// by j descending, the traversal's order) without traversing again, and traverses only
// for bodies whose hits did not fit (a work-group copies the tree only if one of its
// bodies does).
template <bool lds>
__global__ void __launch_bounds__(kBlock) k_bp_step(Dev d, Work w, int pass, int nhits) {
    extern __shared__ __align__(16) unsigned char bp_lds[];
    const int s = blockIdx.x;
    const int b0 = d.body_off[s], nb = d.body_off[s + 1] - b0;
    const int li = blockIdx.y * blockDim.x + threadIdx.x;
    if (blockIdx.y * blockDim.x >= nb) return;  // the whole work-group: no body of this chunk
    const int m = w.nleaves[s];
    bool cached = false;  // pass 1: this body's hits are in its cache slot
    if (pass && li < nb) cached = w.bcnt[b0 + li] + w.bcnt[d.N + b0 + li] <= kBpCache;
    const bool trav = !pass || (li < nb && d.nv[b0 + li] > 0 && !cached);
    const int any_trav = __syncthreads_or(trav);
    BvhLds T{};
    if (lds && any_trav) {
        T.lbox = reinterpret_cast<double*>(bp_lds);
        T.nbox = T.lbox + 6 * m;
        T.child = reinterpret_cast<int*>(T.nbox + 6 * (m > 1 ? m - 1 : 0));
        T.lbody = T.child + 2 * (m > 1 ? m - 1 : 0);
        T.lbid = T.lbody + m;
        T.stk = reinterpret_cast<unsigned short*>(bp_lds + size_t(d.max_nb) * kBvhLdsPerLeaf);
        T.hit = T.stk + kBvhStackBytes / 2;
        for (int k = threadIdx.x; k < m; k += blockDim.x) {
            const int j = w.leaf_body[b0 + k];
            T.lbody[k] = j;
            T.lbid[k] = d.body_id[j];
            for (int q = 0; q < 6; ++q) T.lbox[6 * k + q] = w.wbox[6 * j + q];
        }
        for (int k = threadIdx.x; k < m - 1; k += blockDim.x) {
            T.child[2 * k] = w.node_child[2 * (b0 + k)];
            T.child[2 * k + 1] = w.node_child[2 * (b0 + k) + 1];
            for (int q = 0; q < 6; ++q) T.nbox[6 * k + q] = w.node_box[6 * (b0 + k) + q];
        }
        __syncthreads();
    }
    if (li >= nb) return;
    const int i = b0 + li;
    SceneParams& sp = d.sp[s];
    int* cache = w.bp_cache + size_t(i) * kBpCache;
    if (cached) {
        const int nc = w.bcnt[i], np = w.bcnt[d.N + i];
        const int oc = w.pair_off[s] + w.boff[i], op = w.prox_off[s] + w.boff[d.N + i];
        for (int k = 0; k < nc; ++k) {
            bp_sort_insert(w.pairs, oc, k, i, cache[k]);
            w.pair_scene[oc + k] = s;
        }
        for (int k = 0; k < np; ++k) {
            bp_sort_insert(w.prox_pairs, op, k, i, cache[kBpCache - 1 - k]);
            w.prox_scene[op + k] = s;
        }
        return;
    }
    int nc = 0, np = 0;
    if (d.nv[i] > 0) {
        const double r0 = dmax(0.0, sp.dhat_cand);
        const double r1 = dmax(0.0, sp.dhat * 2.0);
        const double r2 = dmax(0.0, sp.ccd_infl);
        const bool do1 = sp.enable_adaptive != 0, do2 = sp.enable_ccd != 0;
        double rmax = r0;
        if (do1) rmax = dmax(rmax, r1);
        if (do2) rmax = dmax(rmax, r2);
        const Box3 bi = ldbox(w.wbox, i);
        const Box3 q0 = box_inflated(bi, r0), q1 = box_inflated(bi, r1), q2 = box_inflated(bi, r2);
        double gmin = 1.0 / 0.0;
        int any = 0;
        const int oc = pass ? w.boff[i] : 0, op = pass ? w.boff[d.N + i] : 0;
        const double margin = sp.prune_margin;
        const int idi = d.body_id[i];
        // pass 1: hits arriving in descending j (sorted batches) are appended; a hit
        // above the last one written goes to its place by insertion
        int lastc = 0x7fffffff, lastp = 0x7fffffff;
        auto hit = [&](int j, int idj, const Box3& bj) {
            const double gap = box_distance(bi, bj);
            if (box_overlap(q0, box_inflated(bj, r0)) && gap <= r0) {
                if (pass) {
                    const int o = w.pair_off[s] + oc;
                    if (j < lastc) {
                        w.pairs[2 * (o + nc)] = i;
                        w.pairs[2 * (o + nc) + 1] = j;
                        lastc = j;
                    } else {
                        bp_sort_insert(w.pairs, o, nc, i, j);
                    }
                    w.pair_scene[o + nc] = s;
                } else if (nc + np < kBpCache) {
                    cache[nc] = j;
                }
                ++nc;
            }
            if (do1 && box_overlap(q1, box_inflated(bj, r1))) {
                gmin = dmin(gmin, gap);
                if (gap <= margin) {
                    if (pass) {
                        const int o = w.prox_off[s] + op;
                        if (j < lastp) {
                            w.prox_pairs[2 * (o + np)] = i;
                            w.prox_pairs[2 * (o + np) + 1] = j;
                            lastp = j;
                        } else {
                            bp_sort_insert(w.prox_pairs, o, np, i, j);
                        }
                        w.prox_scene[o + np] = s;
                    } else if (nc + np < kBpCache) {
                        cache[kBpCache - 1 - np] = j;
                    }
                    ++np;
                }
            }
            if (do2 && !any && box_overlap(q2, box_inflated(bj, r2)) && pair_allowed(d, s, idi, idj)) any = 1;
        };
        if constexpr (lds)
            lbvh_query_lds(T, m, i, bi, rmax, pass != 0, nhits,
                           [&](int leaf) { hit(T.lbody[leaf], T.lbid[leaf], box_at(T.lbox, leaf)); });
        else
            lbvh_query(d, w, s, i, rmax, [&](int j) { hit(j, d.body_id[j], ldbox(w.wbox, j)); });
        if (!pass) {
            if (do1 && gmin < 1.0 / 0.0)
                atomicMin(reinterpret_cast<unsigned long long*>(&sp.min_pair_dist), __double_as_longlong(gmin));
            if (any) atomicOr(&sp.ccd_any, 1);
        }
    }
    if (!pass) {
        w.bcnt[i] = nc;
        w.bcnt[d.N + i] = np;
    }
}

// Exclusive scans of the per-body counts (candidates, prox) -> per-body offsets within
// their scene and per-scene totals; one work-group per (scene, list). The per-scene
// offsets are then a scan over scenes (k_scan_counts).
__global__ void __launch_bounds__(kBlock) k_bp_scan(Dev d, Work w) {
    __shared__ int sh[kBlock];
    const int s = blockIdx.x, list = blockIdx.y;
    const int* cnt = w.bcnt + list * d.N;
    int* boff = w.boff + list * d.N;
    int* scnt = list == 0 ? w.pair_cnt : w.prox_cnt;
    const int b0 = d.body_off[s], b1 = d.body_off[s + 1];
    int local = 0;
    for (int base = b0; base < b1; base += kBlock) {
        const int b = base + threadIdx.x;
        const int v = b < b1 ? cnt[b] : 0;
        int ex;
        const int t = wg_exclusive_scan(v, sh, &ex);
        if (b < b1) boff[b] = local + ex;
        local += t;
    }
    if (threadIdx.x == 0) scnt[s] = local;
}

__global__ void __launch_bounds__(kBlock) k_scan_counts(const int* cnt, int* off, int* total, int S);

// Dynamic LDS that k_bp_step may request: raised from the 64 KiB default by
// geom_configure (called at context creation, outside any stream capture).
static size_t g_bp_lds_cap = 65536;

int geom_configure() {
    // NDIPC_BP_NO_LDS=1 keeps every traversal on the tree in HBM (the path of scenes too
    // large for LDS), so that tests can run it on small scenes
    const char* e = getenv("NDIPC_BP_NO_LDS");
    if (e && e[0] == '1') {
        g_bp_lds_cap = 0;
        return 0;
    }
    const size_t want = 131072;
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(k_bp_step<true>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, int(want)) == hipSuccess)
        g_bp_lds_cap = want;
    else
        (void)hipGetLastError();
    return 0;
}

void launch_broad_phase_step(const Dev& d, const Work& w, hipStream_t st) {
    if (d.N == 0) return;
    const dim3 grid(d.S, (d.max_nb + kBlock - 1) / kBlock);
    // the tree of the largest scene in LDS, with the lanes' traversal stacks and hit
    // buffers (the rest of the LDS, at least kBvhMinHits entries a lane: 548 bodies with
    // 128 KiB), otherwise the tree in HBM. A lane whose hits overflow its buffer emits
    // them batch by batch (later batches by insertion).
    const size_t fixed = size_t(d.max_nb) * kBvhLdsPerLeaf + kBvhStackBytes;
    int nhits = 0;
    if (fixed < g_bp_lds_cap) nhits = int(std::min<size_t>((g_bp_lds_cap - fixed) / (2 * kBlock), 1024));
    const int lds = nhits >= kBvhMinHits ? 1 : 0;
    const size_t bytes = fixed + size_t(nhits) * 2 * kBlock;
    timing_begin("bp_step", st);
    if (lds) hipLaunchKernelGGL(k_bp_step<true>, grid, dim3(kBlock), bytes, st, d, w, 0, nhits);
    else hipLaunchKernelGGL(k_bp_step<false>, grid, dim3(kBlock), 0, st, d, w, 0, 0);
    timing_end("bp_step", st);
    hipLaunchKernelGGL(k_bp_scan, dim3(d.S, 2), dim3(kBlock), 0, st, d, w);
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(kBlock), 0, st, w.pair_cnt, w.pair_off, w.n_pairs_total, d.S);
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(kBlock), 0, st, w.prox_cnt, w.prox_off, w.n_prox_total, d.S);
    timing_begin("bp_step", st);
    if (lds) hipLaunchKernelGGL(k_bp_step<true>, grid, dim3(kBlock), bytes, st, d, w, 1, nhits);
    else hipLaunchKernelGGL(k_bp_step<false>, grid, dim3(kBlock), 0, st, d, w, 1, 0);
    timing_end("bp_step", st);
}

// mode 0: candidate pairs (gap <= r); mode 3: all overlapping pairs (detectBodyPairs).
// mode 1: min gap over pairs at 2*dhat + prox pairs (gap <= prune_margin).
// mode 2: CCD gate — any allowed pair at inflation max(1e-6, dhat) + vmax dt.
// pass 0 counts, pass 1 emits at pair_off[s] (scan of the counts).
__global__ void __launch_bounds__(kBlock) k_broad_phase(Dev d, Work w, int mode, int pass, double r_in,
                                                        int only_scene) {
    __shared__ int shi[kBlock];
    __shared__ double shd[kBlock];
    const int s = blockIdx.x;
    const SceneParams& sp = d.sp[s];
    const int t = threadIdx.x;
    const int b0 = d.body_off[s], b1 = d.body_off[s + 1];
    int* cnt = (mode == 1) ? w.prox_cnt : w.pair_cnt;
    const bool step_mode = r_in < 0.0;  // per-scene inflation of IPCSolver::step
    const bool skip = (only_scene >= 0 && s != only_scene) || (step_mode && mode == 1 && !sp.enable_adaptive) ||
                      (step_mode && mode == 2 && !sp.enable_ccd);
    if (skip) {
        if (pass == 0 && t == 0 && mode != 2) cnt[s] = 0;
        if (mode == 1 && pass == 0 && t == 0) d.sp[s].min_pair_dist = 1.0 / 0.0;
        if (mode == 2 && t == 0) d.sp[s].ccd_any = 0;
        return;
    }
    double r = r_in;
    if (step_mode) r = (mode == 1) ? sp.dhat * 2.0 : (mode == 2 ? sp.ccd_infl : sp.dhat);
    r = dmax(0.0, r);
    int* off = (mode == 1) ? w.prox_off : w.pair_off;
    int* outp = (mode == 1) ? w.prox_pairs : w.pairs;
    int* outs = (mode == 1) ? w.prox_scene : w.pair_scene;
    const double margin = d.sp[s].prune_margin;
    int running = 0;
    double gap_min = 1.0 / 0.0;
    int any = 0;
    for (int base = b0; base < b1; base += kBlock) {
        const int i = base + t;
        int c = 0;
        if (i < b1 && d.nv[i] > 0) {
            lbvh_query(d, w, s, i, r, [&](int j) {
                const double gap = box_distance(ldbox(w.wbox, i), ldbox(w.wbox, j));
                if (mode == 0) { if (gap <= r) ++c; }
                else if (mode == 3) ++c;
                else if (mode == 1) { gap_min = dmin(gap_min, gap); if (gap <= margin) ++c; }
                else if (pair_allowed(d, s, d.body_id[i], d.body_id[j])) any = 1;
            });
        }
        if (mode == 2) continue;
        int ex;
        const int tot = wg_exclusive_scan(c, shi, &ex);
        if (pass == 1 && c > 0) {
            const int o = off[s] + running + ex;
            int k = 0;
            lbvh_query(d, w, s, i, r, [&](int j) {
                const double gap = box_distance(ldbox(w.wbox, i), ldbox(w.wbox, j));
                const bool keep = (mode == 0) ? (gap <= r) : (mode == 3 ? true : (gap <= margin));
                if (keep && k < c) {
                    // insertion into the segment in the reference's order (j descending)
                    int pos = k;
                    while (pos > 0 && outp[2 * (o + pos - 1) + 1] < j) {
                        outp[2 * (o + pos)] = outp[2 * (o + pos - 1)];
                        outp[2 * (o + pos) + 1] = outp[2 * (o + pos - 1) + 1];
                        --pos;
                    }
                    outp[2 * (o + pos)] = i;
                    outp[2 * (o + pos) + 1] = j;
                    outs[o + k] = s;
                    ++k;
                }
            });
        }
        running += tot;
    }
    if (mode == 2) {
        const int a = wg_reduce(any, shi, OpOrI());
        if (t == 0) d.sp[s].ccd_any = a;
        return;
    }
    if (pass == 0 && t == 0) cnt[s] = running;
    if (mode == 1 && pass == 0) {
        const double g = wg_reduce(gap_min, shd, OpMin());
        if (t == 0) d.sp[s].min_pair_dist = g;
    }
}

// Exclusive scan of per-scene counts (one work-group).
__global__ void __launch_bounds__(kBlock) k_scan_counts(const int* cnt, int* off, int* total, int S) {
    __shared__ int sh[kBlock];
    int running = 0;
    for (int base = 0; base < S; base += kBlock) {
        const int i = base + threadIdx.x;
        const int v = i < S ? cnt[i] : 0;
        int ex;
        const int tot = wg_exclusive_scan(v, sh, &ex);
        if (i < S) off[i] = running + ex;
        running += tot;
    }
    if (threadIdx.x == 0) {
        off[S] = running;
        *total = running;
    }
}

void launch_broad_phase(const Dev& d, const Work& w, int mode, double r, bool, const int* scene_mask,
                        hipStream_t st) {
    const int only = scene_mask ? *scene_mask : -1;
    hipLaunchKernelGGL(k_broad_phase, dim3(d.S), dim3(kBlock), 0, st, d, w, mode, 0, r, only);
    if (mode == 2) return;
    int* cnt = (mode == 1) ? w.prox_cnt : w.pair_cnt;
    int* off = (mode == 1) ? w.prox_off : w.pair_off;
    int* tot = (mode == 1) ? w.n_prox_total : w.n_pairs_total;
    if (only >= 0) {
        // single-scene call: other scenes' counts are zeroed on the host side of the API
    }
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(kBlock), 0, st, cnt, off, tot, d.S);
    hipLaunchKernelGGL(k_broad_phase, dim3(d.S), dim3(kBlock), 0, st, d, w, mode, 1, r, only);
}

// ------------------------------------------------------------------------ top-K
// select_top_k_indices (ContactAssembler.cpp:67-88) for the six lists of every close
// pair (ContactAssembler.cpp:249-302). Scores are point_aabb_distance of world
// vertices to the other body's world AABB; edges take the min over 2 endpoints, faces
// over 3 (std::min / std::min({..}) return the smallest value). The kept set and its
// order are those of std::nth_element followed by std::sort (libstdc++): with equal
// scores among the kept ones (most lists: an edge or face score is the score of a
// shared vertex, and vertices inside the other box all score 0) both depend on the
// library's algorithms, so every list replays them. One wave (a 64-lane work-group)
// owns a list, held in LDS as (score, 16-bit index) pairs, and runs the steps of
// tk_nth_element_lr and tk_sort_blocks (topk_select.h, checked against libstdc++ on
// the host): each Hoare partition is computed from the positions it swaps with a wave
// scan, median-of-three and the heap fallbacks run on lane 0, and the final insertion
// sorts are per-block rank sorts with one lane per element.
// Lists of up to kTkLdsPrims entries live in the wave's LDS; longer ones (meshes of up to
// kMaxPrims primitives) run the same steps on the same arrays in a per-work-group slab
// of HBM (TkHbm), with only the range stack in LDS.
__device__ __forceinline__ bool key_less(double sa, int ia, double sb, int ib) {
    return sa < sb || (sa == sb && ia < ib);
}

constexpr int kTkLanes = 64;
constexpr int kTkPer = kTkLdsPrims / kTkLanes;  // kept entries per lane at most (k <= kTkLdsPrims)

struct TkLds {
    double key[kTkLdsPrims];
    unsigned short idx[kTkLdsPrims];
    unsigned short lpos[kTkLdsPrims], rpos[kTkLdsPrims];
    unsigned char start[kTkLdsPrims];
    int stk[3 * 64];
};

// The arrays of TkLds for a list of up to M <= kMaxPrims entries in a slab of HBM
// (tk_slab_bytes(M) bytes, 256-byte aligned parts); the range stack stays in LDS.
struct TkHbm {
    double* key;
    unsigned short *idx, *lpos, *rpos;
    unsigned char* start;
    int* stk;
};

__host__ __device__ inline size_t tk_part(size_t bytes) { return (bytes + 255) & ~size_t(255); }
size_t tk_slab_bytes(int M) {
    return tk_part(8 * size_t(M)) + 3 * tk_part(2 * size_t(M)) + tk_part(size_t(M));
}
__device__ inline TkHbm tk_hbm(unsigned char* base, int M, int* stk) {
    TkHbm h;
    h.key = reinterpret_cast<double*>(base);
    base += tk_part(8 * size_t(M));
    h.idx = reinterpret_cast<unsigned short*>(base);
    base += tk_part(2 * size_t(M));
    h.lpos = reinterpret_cast<unsigned short*>(base);
    base += tk_part(2 * size_t(M));
    h.rpos = reinterpret_cast<unsigned short*>(base);
    base += tk_part(2 * size_t(M));
    h.start = base;
    h.stk = stk;
    return h;
}

// Exclusive prefix sum over the wave's 64 lanes, and the total (unsigned: two 16-bit
// counts of lists up to 65535 entries are packed in one word).
__device__ __forceinline__ unsigned tk_wave_excl_sum(unsigned v, unsigned* total) {
    unsigned x = v;
#pragma unroll
    for (int o = 1; o < kTkLanes; o <<= 1) {
        const unsigned y = __shfl_up(x, o, kTkLanes);
        if ((int)threadIdx.x >= o) x += y;
    }
    *total = __shfl(x, kTkLanes - 1, kTkLanes);
    return x - v;
}

// tk_partition_lr over [lo, hi) with the pivot at `pivot`: L (ascending positions not
// below the pivot) and R (descending positions not above it) by a wave scan of per-lane
// counts over contiguous chunks; m = the first q with L_q >= R_q (L ascends and R
// descends, so the test is monotone in q: a 64-way search in two rounds); the m swaps
// touch 2m distinct positions and run in parallel. Returns the cut (wave-uniform).
__device__ int tk_wave_partition(const KV16Store& a, int lo, int hi, int pivot, unsigned short* lpos,
                                 unsigned short* rpos) {
    const int lane = threadIdx.x;
    const double pv = a.k[pivot];
    const int len = hi - lo;
    const int per = (len + kTkLanes - 1) / kTkLanes;
    const int i0 = lo + lane * per, i1 = min(hi, i0 + per);
    unsigned cl = 0, cr = 0;
    for (int i = i0; i < i1; ++i) {
        const double v = a.k[i];
        cl += !(v < pv);
        cr += !(pv < v);
    }
    unsigned tot;
    const unsigned ex = tk_wave_excl_sum(cl | (cr << 16), &tot);
    const int nL = int(tot & 0xffffu), nR = int(tot >> 16);
    int kl = int(ex & 0xffffu), kr = int(ex >> 16);
    for (int i = i0; i < i1; ++i) {
        const double v = a.k[i];
        if (!(v < pv)) lpos[kl++] = (unsigned short)i;
        if (!(pv < v)) rpos[nR - 1 - kr++] = (unsigned short)i;
    }
    __syncthreads();
    // m by tk_first_crossing: 64 probes a step apart, then the probes of the one step
    // that holds the crossing, 64 at a time (step > 64 for lists above 4096 entries)
    const int mm = min(nL, nR);
    const int step = (mm + kTkLanes - 1) / kTkLanes;
    const int nq = step ? (mm + step - 1) / step : 0;  // probes below mm
    int m;
    {
        const int q = lane * step;
        const unsigned long long b = __ballot(lane < nq && lpos[q] >= rpos[q]);
        const int f = b ? __ffsll((unsigned long long)b) - 1 : nq;
        if (f == 0) {
            m = 0;
        } else {  // m in ((f - 1) * step, min(f * step, mm)]
            const int qa = (f - 1) * step + 1, qb = min(f * step, mm);
            m = qb;
            for (int base = qa; base < qb; base += kTkLanes) {
                const int q2 = base + lane;
                const unsigned long long b2 = __ballot(q2 < qb && lpos[q2] >= rpos[q2]);
                if (b2) {
                    m = base + __ffsll((unsigned long long)b2) - 1;
                    break;
                }
            }
        }
    }
    int cut = m < nL ? (int)lpos[m] : hi;
    if (m >= 1 && (int)rpos[m - 1] < cut) cut = rpos[m - 1];
    for (int q = lane; q < m; q += kTkLanes) {
        const int x = lpos[q], y = rpos[q];
        const double tk = a.k[x];
        const unsigned short ti = a.idx[x];
        a.k[x] = a.k[y];
        a.idx[x] = a.idx[y];
        a.k[y] = tk;
        a.idx[y] = ti;
    }
    __syncthreads();
    return cut;
}

// tk_block_rank_sort of [0, n), n <= 64 * PER: lane-parallel, every element's
// destination computed (and the element held in registers) before any element moves.
template <int PER>
__device__ void tk_wave_block_rank_sort(const KV16Store& a, int n, const unsigned char* start) {
    const int lane = threadIdx.x;
    int dst[PER];
    double kv[PER];
    unsigned short iv[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int i = lane + j * kTkLanes;
        if (i < n) {
            int s = i, e = i + 1;
            while (!start[s]) --s;
            while (e < n && !start[e]) ++e;
            const double v = a.k[i];
            int r = s;
            for (int t = s; t < e; ++t) {
                const double u = a.k[t];
                r += (u < v) || (t < i && !(v < u));
            }
            dst[j] = r;
            kv[j] = v;
            iv[j] = a.idx[i];
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int i = lane + j * kTkLanes;
        if (i < n) {
            a.k[dst[j]] = kv[j];
            a.idx[dst[j]] = iv[j];
        }
    }
    __syncthreads();
}

// tk_nth_element_lr(a, nth, n) by the wave.
template <class TL>
__device__ void tk_wave_nth_element(const KV16Store& a, int nth, int n, TL& L) {
    if (n == 0 || nth == n) return;
    const int lane = threadIdx.x;
    int first = 0, last = n;
    int depth = tk_floor_log2(n) * 2;
    while (last - first > 3) {
        if (depth == 0) {
            if (lane == 0) {
                tk_heap_select(a, first, nth + 1, last);
                tk_swap(a, first, nth);
            }
            __syncthreads();
            return;
        }
        --depth;
        if (lane == 0) tk_move_median_to_first(a, first, first + 1, first + (last - first) / 2, last - 1);
        __syncthreads();
        const int cut = tk_wave_partition(a, first + 1, last, first, L.lpos, L.rpos);
        if (cut <= nth) first = cut;
        else last = cut;
    }
    // the final insertion sort of at most 3 elements: a one-block rank sort
    const int len = last - first;
    double v = 0.0;
    unsigned short vi = 0;
    int r = 0;
    if (lane < len) {
        v = a.k[first + lane];
        vi = a.idx[first + lane];
        for (int t = 0; t < len; ++t) {
            const double u = a.k[first + t];
            r += (u < v) || (t < lane && !(v < u));
        }
    }
    __syncthreads();
    if (lane < len) {
        a.k[first + r] = v;
        a.idx[first + r] = vi;
    }
    __syncthreads();
}

// The introsort loop of tk_sort_blocks(a, n) by the wave: leaves [0, n) cut into blocks
// (L.start[i] = 1 where one begins) that its final insertion sort orders one by one.
// Lane 0 keeps the range stack in LDS.
template <class TL>
__device__ void tk_wave_sort_loop(const KV16Store& a, int n, TL& L) {
    const int lane = threadIdx.x;
    int* sf = L.stk;
    int* sl = L.stk + 64;
    int* sd = L.stk + 128;
    for (int i = lane; i < n; i += kTkLanes) L.start[i] = 0;
    if (lane == 0) {
        sf[0] = 0;
        sl[0] = n;
        sd[0] = 2 * tk_floor_log2(n);
    }
    int sp = 1;
    while (sp > 0) {
        __syncthreads();
        --sp;
        const int first = sf[sp];
        int last = sl[sp], depth = sd[sp];
        bool heap = false;
        while (last - first > 16) {
            if (depth == 0) {
                if (lane == 0) {
                    tk_heap_select(a, first, last, last);
                    tk_sort_heap(a, first, last);
                }
                for (int i = first + lane; i < last; i += kTkLanes) L.start[i] = 1;
                heap = true;
                break;
            }
            --depth;
            if (lane == 0) tk_move_median_to_first(a, first, first + 1, first + (last - first) / 2, last - 1);
            __syncthreads();
            const int cut = tk_wave_partition(a, first + 1, last, first, L.lpos, L.rpos);
            if (lane == 0) {
                sf[sp] = cut;
                sl[sp] = last;
                sd[sp] = depth;
            }
            ++sp;
            last = cut;
        }
        if (!heap && lane == 0 && first < n) L.start[first] = 1;
    }
    __syncthreads();
}

// tk_sort_blocks(a, n) by the wave, n <= 64 * PER.
template <int PER, class TL>
__device__ void tk_wave_sort(const KV16Store& a, int n, TL& L) {
    if (n < 2) return;
    tk_wave_sort_loop(a, n, L);
    tk_wave_block_rank_sort<PER>(a, n, L.start);
}

// The list's scores are in L.key[0..n) with L.idx[i] = i; leaves the kept k indices
// (k <= 64 * PER), in the reference's order, in L.idx[0..k).
template <int PER, class TL>
__device__ void tk_wave_select(TL& L, int n, int k) {
    const KV16Store a{L.key, L.idx};
    if (k < n) tk_wave_nth_element(a, k, n, L);
    tk_wave_sort<PER>(a, k, L);
}

// Scores of list `list` of close pair p into L.key[0..n) (L.idx[i] = i), the selection,
// and the kept indices out to w.lists.
template <class TL>
__device__ void tk_list(const Dev& d, const Work& w, TL& L, int p, int list, int n, int X, int Y) {
    const int lane = threadIdx.x;
    const Box3 by = ldbox(w.wbox, Y);
    const V3 px = ld3(d.p, X);
    const Q qx = ld4(d.q, X);
    const int kind = list >> 1;  // 0 vertices, 1 edges, 2 faces
    for (int i = lane; i < n; i += kTkLanes) {
        double sc;
        if (kind == 0) {
            sc = point_box_distance(world_vertex(d, X, i, px, qx), by);
        } else if (kind == 1) {
            const int* e = d.edges + 2 * (d.e_off[X] + i);
            const double a = point_box_distance(world_vertex(d, X, e[0], px, qx), by);
            const double b = point_box_distance(world_vertex(d, X, e[1], px, qx), by);
            sc = (b < a) ? b : a;
        } else {
            const int* f = d.faces + 3 * (d.f_off[X] + i);
            const double a = point_box_distance(world_vertex(d, X, f[0], px, qx), by);
            const double b = point_box_distance(world_vertex(d, X, f[1], px, qx), by);
            const double c = point_box_distance(world_vertex(d, X, f[2], px, qx), by);
            sc = a;
            if (b < sc) sc = b;
            if (c < sc) sc = c;
        }
        L.key[i] = sc;
        L.idx[i] = (unsigned short)i;
    }
    __syncthreads();
    const int k = n < kTopK ? n : kTopK;
    tk_wave_select<(kTopK + kTkLanes - 1) / kTkLanes>(L, n, k);
    for (int i = lane; i < k; i += kTkLanes) w.lists[(p * 6 + list) * kTopK + i] = L.idx[i];
    if (lane == 0) w.lcnt[p * 6 + list] = k;
    __syncthreads();
}

// One wave per list. A list longer than kTkLdsPrims uses work-group blockIdx.x's slab of
// w.tk_slab (allocated at upload for such meshes; launch_topk keeps the grid within
// w.tk_slab_blocks then).
__global__ void __launch_bounds__(kTkLanes) k_topk(Dev d, Work w) {
    __shared__ TkLds L;
    const int total = *w.n_pairs_total;
    for (int task = blockIdx.x; task < total * 6; task += gridDim.x) {
        const int p = task / 6, list = task % 6;
        const int A = w.pairs[2 * p], B = w.pairs[2 * p + 1];
        const int X = (list & 1) ? B : A;
        const int Y = (list & 1) ? A : B;
        const int kind = list >> 1;
        const int n = kind == 0 ? d.nv[X] : (kind == 1 ? d.ne[X] : d.nf[X]);
        if (n <= kTkLdsPrims) {
            tk_list(d, w, L, p, list, n, X, Y);
        } else {
            TkHbm H = tk_hbm(w.tk_slab + size_t(blockIdx.x) * w.tk_slab_stride, w.tk_slab_m, L.stk);
            tk_list(d, w, H, p, list, n, X, Y);
        }
    }
}

void launch_topk(const Dev& d, const Work& w, int max_pairs, hipStream_t st) {
    int cap = w.grid_topk > 0 ? w.grid_topk : 4096;
    if (w.tk_slab && cap > w.tk_slab_blocks) cap = w.tk_slab_blocks;
    int grid = max_pairs * 6;
    if (grid > cap) grid = cap;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(k_topk, dim3(grid), dim3(kTkLanes), 0, st, d, w);
}

template <class TL>
__device__ void tk_select_given(TL& L, const double* scores, int n, int kk, int* out, int* out_n) {
    const int lane = threadIdx.x;
    for (int i = lane; i < n; i += kTkLanes) {
        L.key[i] = scores[i];
        L.idx[i] = (unsigned short)i;
    }
    __syncthreads();
    tk_wave_select<kTkPer>(L, n, kk);
    for (int i = lane; i < kk; i += kTkLanes) out[i] = L.idx[i];
    if (lane == 0) *out_n = kk;
}

// Standalone selection on given scores (one wave): n <= kTkLdsPrims in LDS, else in
// `slab` (tk_slab_bytes(n) bytes); k <= kTkLdsPrims.
__global__ void __launch_bounds__(kTkLanes) k_select_scores(const double* scores, int n, int k, int* out,
                                                           int* out_n, unsigned char* slab) {
    __shared__ TkLds L;
    const int kk = k < n ? k : n;
    if (kk <= 0) {
        if (threadIdx.x == 0) *out_n = 0;
        return;
    }
    if (n <= kTkLdsPrims) {
        tk_select_given(L, scores, n, kk, out, out_n);
    } else {
        TkHbm H = tk_hbm(slab, n, L.stk);
        tk_select_given(H, scores, n, kk, out, out_n);
    }
}

void launch_select_topk_scores(const double* scores, int n, int k, int* out, int* out_n, unsigned char* slab,
                               hipStream_t st) {
    hipLaunchKernelGGL(k_select_scores, dim3(1), dim3(kTkLanes), 0, st, scores, n, k, out, out_n, slab);
}

// ------------------------------------------------------------------------- poses
// Body poses of the narrow phase: step-start state (x == NULL, detectContacts) or the
// Newton iterate (detectContactsAtState, IPCSolver.cpp:604-629).
__global__ void k_pose(Dev d, Work w, const double* x, int gate) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= d.N) return;
    if (gate && d.sp[d.body_scene[b]].alm_done) return;
    if (!x && as_skip(w, d.sp[d.body_scene[b]])) return;
    if (!x) {
        st3(w.pose_p, b, ld3(d.p, b));
        st4(w.pose_q, b, ld4(d.q, b));
    } else {
        st3(w.pose_p, b, v3(x[6 * b], x[6 * b + 1], x[6 * b + 2]));
        st4(w.pose_q, b, apply_theta(ld4(d.q, b), v3(x[6 * b + 3], x[6 * b + 4], x[6 * b + 5])));
    }
}

void launch_pose_from_state(const Dev& d, const Work& w, const double* x, hipStream_t st) {
    if (d.N == 0) return;
    hipLaunchKernelGGL(k_pose, dim3((d.N + kBlock - 1) / kBlock), dim3(kBlock), 0, st, d, w, x, x ? 1 : 0);
}

// ------------------------------------------------------------------ narrow phase
// IPCSolver::detectContacts / detectContactsAtState (IPCSolver.cpp:394-806).
// Candidate order per pair: FV faces(A) x verts(B), FV faces(B) x verts(A), then
// EE edges(A) x edges(B) (ContactAssembler.cpp:304-341); all FV hits of a scene come
// before its EE hits. Bits: [0,1024) FV A-B, [1024,2048) FV B-A, [2048,3072) EE.

struct PairGeom {
    V3 vA[kTopK], vB[kTopK];
    V3 fA[kTopK][3], fB[kTopK][3];
    V3 nA[kTopK], nB[kTopK];  // safe-normalized face normals
    V3 eA[kTopK][2], eB[kTopK][2];
    int lvA[kTopK], lvB[kTopK], lfA[kTopK], lfB[kTopK], leA[kTopK][2], leB[kTopK][2];
    int k[6];
    int A, B, s, allowed;
    double dhat;
};

// safeNormalized (IPCSolver.cpp:92-105)
__device__ __forceinline__ V3 safe_normalized(V3 v, V3 fb) {
    const double n = norm(v);
    if (n > 1e-12) return v / n;
    const double fn = norm(fb);
    if (fn > 1e-12) return fb / fn;
    return v3(0.0, 1.0, 0.0);
}

// pointInTriangleBarycentric (IPCSolver.cpp:107-134)
__device__ __forceinline__ bool point_in_triangle(V3 p, V3 a, V3 b, V3 c) {
    const double eps = 1e-8;
    const V3 v0 = b - a, v1 = c - a, v2 = p - a;
    const double d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1);
    const double d20 = dot(v2, v0), d21 = dot(v2, v1);
    const double den = d00 * d11 - d01 * d01;
    if (fabs(den) < 1e-16) return false;
    const double v = (d11 * d20 - d01 * d21) / den;
    const double ww = (d00 * d21 - d01 * d20) / den;
    const double u = 1.0 - v - ww;
    return u >= -eps && v >= -eps && ww >= -eps;
}

__device__ __forceinline__ double clamp01(double x) { return x < 0.0 ? 0.0 : (1.0 < x ? 1.0 : x); }

// closestPointsOnSegments (IPCSolver.cpp:32-90)
__device__ __forceinline__ void closest_points(V3 p1, V3 q1, V3 p2, V3 q2, V3& c1, V3& c2) {
    const double eps = 1e-12;
    const V3 d1 = q1 - p1, d2 = q2 - p2, r = p1 - p2;
    const double a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    double s, t;
    if (a <= eps && e <= eps) { c1 = p1; c2 = p2; return; }
    if (a <= eps) {
        s = 0.0;
        t = clamp01(f / e);
    } else {
        const double c = dot(d1, r);
        if (e <= eps) {
            t = 0.0;
            s = clamp01(-c / a);
        } else {
            const double b = dot(d1, d2);
            const double den = a * e - b * b;
            s = (fabs(den) > eps) ? clamp01((b * f - c * e) / den) : 0.0;
            t = (b * s + f) / e;
            if (t < 0.0) { t = 0.0; s = clamp01(-c / a); }
            else if (t > 1.0) { t = 1.0; s = clamp01((b - c) / a); }
        }
    }
    c1 = p1 + d1 * s;
    c2 = p2 + d2 * t;
}

__device__ void load_pair_geom(const Dev& d, const Work& w, int p, PairGeom& g) {
    const int t = threadIdx.x;
    if (t == 0) {
        g.A = w.pairs[2 * p];
        g.B = w.pairs[2 * p + 1];
        g.s = w.pair_scene[p];
        for (int l = 0; l < 6; ++l) g.k[l] = w.lcnt[p * 6 + l];
        g.allowed = pair_allowed(d, g.s, d.body_id[g.A], d.body_id[g.B]) ? 1 : 0;
        g.dhat = d.sp[g.s].dhat;
    }
    __syncthreads();
    const int A = g.A, B = g.B;
    const V3 pA = ld3(w.pose_p, A), pB = ld3(w.pose_p, B);
    const Q qA = ld4(w.pose_q, A), qB = ld4(w.pose_q, B);
    const int* L = w.lists + p * 6 * kTopK;
    if (t < kTopK) {
        const int i = t;
        if (i < g.k[0]) { g.lvA[i] = L[0 * kTopK + i]; g.vA[i] = world_vertex(d, A, g.lvA[i], pA, qA); }
        if (i < g.k[1]) { g.lvB[i] = L[1 * kTopK + i]; g.vB[i] = world_vertex(d, B, g.lvB[i], pB, qB); }
    } else if (t < 2 * kTopK) {
        const int i = t - kTopK;
        if (i < g.k[2]) {
            const int e = L[2 * kTopK + i];
            const int* ev = d.edges + 2 * (d.e_off[A] + e);
            g.leA[i][0] = ev[0]; g.leA[i][1] = ev[1];
            g.eA[i][0] = world_vertex(d, A, ev[0], pA, qA);
            g.eA[i][1] = world_vertex(d, A, ev[1], pA, qA);
        }
        if (i < g.k[3]) {
            const int e = L[3 * kTopK + i];
            const int* ev = d.edges + 2 * (d.e_off[B] + e);
            g.leB[i][0] = ev[0]; g.leB[i][1] = ev[1];
            g.eB[i][0] = world_vertex(d, B, ev[0], pB, qB);
            g.eB[i][1] = world_vertex(d, B, ev[1], pB, qB);
        }
    } else if (t < 3 * kTopK) {
        const int i = t - 2 * kTopK;
        if (i < g.k[4]) {
            const int f = L[4 * kTopK + i];
            g.lfA[i] = f;
            const int* fv = d.faces + 3 * (d.f_off[A] + f);
            for (int c = 0; c < 3; ++c) g.fA[i][c] = world_vertex(d, A, fv[c], pA, qA);
            g.nA[i] = safe_normalized(cross(g.fA[i][1] - g.fA[i][0], g.fA[i][2] - g.fA[i][0]), v3(0.0, 1.0, 0.0));
        }
        if (i < g.k[5]) {
            const int f = L[5 * kTopK + i];
            g.lfB[i] = f;
            const int* fv = d.faces + 3 * (d.f_off[B] + f);
            for (int c = 0; c < 3; ++c) g.fB[i][c] = world_vertex(d, B, fv[c], pB, qB);
            g.nB[i] = safe_normalized(cross(g.fB[i][1] - g.fB[i][0], g.fB[i][2] - g.fB[i][0]), v3(0.0, 1.0, 0.0));
        }
    }
    __syncthreads();
}

// FV test of face f (verts tri, normal n) against vertex p.
__device__ __forceinline__ bool fv_test(const V3* tri, V3 n, V3 p, double dhat, double& sd, V3& proj) {
    sd = dot(p - tri[0], n);
    const double dist = fabs(sd);
    if (dist >= dhat) return false;
    proj = p - sd * n;
    return point_in_triangle(proj, tri[0], tri[1], tri[2]);
}

__device__ __forceinline__ bool cand_hit(const PairGeom& g, int c, bool vv) {
    if (!g.allowed) return false;
    if (vv) {
        const int kb = g.k[1];
        if (c >= g.k[0] * kb) return false;
        const V3 dd = g.vB[c % kb] - g.vA[c / kb];
        return norm(dd) < g.dhat;
    }
    double sd;
    V3 proj;
    if (c < 1024) {
        const int kb = g.k[1];
        if (kb == 0 || c >= g.k[4] * kb) return false;
        const int i = c / kb, j = c % kb;
        return fv_test(g.fA[i], g.nA[i], g.vB[j], g.dhat, sd, proj);
    }
    if (c < 2048) {
        const int cc = c - 1024;
        const int ka = g.k[0];
        if (ka == 0 || cc >= g.k[5] * ka) return false;
        const int i = cc / ka, j = cc % ka;
        return fv_test(g.fB[i], g.nB[i], g.vA[j], g.dhat, sd, proj);
    }
    const int cc = c - 2048;
    const int kb = g.k[3];
    if (kb == 0 || cc >= g.k[2] * kb) return false;
    const int i = cc / kb, j = cc % kb;
    V3 c0, c1;
    closest_points(g.eA[i][0], g.eA[i][1], g.eB[j][0], g.eB[j][1], c0, c1);
    return norm(c1 - c0) < g.dhat;
}

__global__ void __launch_bounds__(kBlock) k_np_count(Dev d, Work w, int vv, int gate) {
    __shared__ PairGeom g;
    __shared__ int cnt[3];
    const int total = *w.n_pairs_total;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int p = blockIdx.x; p < total; p += gridDim.x) {
        const int s = w.pair_scene[p];
        const SceneParams& sp = d.sp[s];
        if (gate ? sp.alm_done : as_skip(w, sp)) continue;
        if (vv && !sp.vv_needed) continue;  // VV only when no FV/EE contact (IPCSolver.cpp:546)
        if (threadIdx.x < 3) cnt[threadIdx.x] = 0;
        load_pair_geom(d, w, p, g);
        const int nwords = vv ? 16 : kHitWords;
        for (int word = wave; word < nwords; word += 4) {
            const int c = word * 64 + lane;
            const bool h = cand_hit(g, c, vv != 0);
            const unsigned long long m = __ballot(h);
            if (lane == 0) {
                w.hits[(long long)p * kHitWords + word] = m;
                const int pc = __popcll(m);
                atomicAdd(&cnt[(vv || c < 2048) ? 0 : 1], pc);
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            w.hit_fv[p] = cnt[0];
            w.hit_ee[p] = vv ? 0 : cnt[1];
        }
        __syncthreads();
    }
}

// Per scene: offsets of FV hits then EE hits in pair order; raw count capped at 40000.
__global__ void __launch_bounds__(kBlock) k_np_scan(Dev d, Work w, int vv, int gate) {
    __shared__ int sh[kBlock];
    __shared__ double shd[kBlock];
    const int s = blockIdx.x;
    SceneParams& sp = d.sp[s];
    if (gate ? sp.alm_done : as_skip(w, sp)) return;
    if (vv && !sp.vv_needed) return;
    const int p0 = w.pair_off[s], p1 = w.pair_off[s + 1];
    int run = 0;
    double ev = 0.0;
    for (int pass = 0; pass < 2; ++pass) {
        for (int base = p0; base < p1; base += kBlock) {
            const int p = base + threadIdx.x;
            int v = 0;
            double e = 0.0;
            if (p < p1) {
                v = pass == 0 ? w.hit_fv[p] : w.hit_ee[p];
                const int* k = w.lcnt + p * 6;
                if (vv) e = pass == 0 ? double(k[0] * k[1]) : 0.0;
                else e = pass == 0 ? double(k[4] * k[1] + k[5] * k[0]) : double(k[2] * k[3]);
            }
            int ex;
            const int tot = wg_exclusive_scan(v, sh, &ex);
            if (p < p1) {
                if (pass == 0) w.hit_off_fv[p] = run + ex;
                else w.hit_off_ee[p] = run + ex;
            }
            run += tot;
            ev += wg_reduce(e, shd, OpAdd());
        }
    }
    if (threadIdx.x == 0) {
        w.raw_cnt[s] = run < kMaxRawContacts ? run : kMaxRawContacts;
        if (!vv) sp.vv_needed = (run == 0) ? 1 : 0;
        sp.evals += ev;
        sp.evals_total += ev;
    }
}

// append_contact (IPCSolver.cpp:431-469 / 651-695): orient n toward pB - pA (with
// safeNormalized fallback), distance = max(0, d).
__device__ void write_contact(const Dev& d, const Work& w, int s, int pos, int A, int B, int type, int pa, int pb,
                              V3 pt, V3 nrm, double dist, int pa2, int pb2) {
    const V3 pAp = ld3(w.pose_p, A), pBp = ld3(w.pose_p, B);
    V3 n = safe_normalized(nrm, pBp - pAp);
    if (dot(n, pBp - pAp) < 0.0) n = -n;
    const long long c = (long long)s * kMaxRawContacts + pos;
    int* ci = w.cint + 5 * c;
    ci[0] = A; ci[1] = B; ci[2] = type; ci[3] = pa; ci[4] = pb;
    w.cext[2 * c] = pa2;
    w.cext[2 * c + 1] = pb2;
    double* cd = w.cdbl + 7 * c;
    cd[0] = pt.x; cd[1] = pt.y; cd[2] = pt.z;
    cd[3] = n.x; cd[4] = n.y; cd[5] = n.z;
    cd[6] = dist > 0.0 ? dist : 0.0;
}

__global__ void __launch_bounds__(kBlock) k_np_write(Dev d, Work w, int vv, int gate) {
    __shared__ PairGeom g;
    __shared__ int wpre[kHitWords];
    const int total = *w.n_pairs_total;
    for (int p = blockIdx.x; p < total; p += gridDim.x) {
        const int s = w.pair_scene[p];
        if (gate ? d.sp[s].alm_done : as_skip(w, d.sp[s])) continue;
        if (vv && (!d.sp[s].vv_needed || w.hit_fv[p] == 0)) continue;
        if (!vv && w.hit_fv[p] + w.hit_ee[p] == 0) continue;
        load_pair_geom(d, w, p, g);
        const unsigned long long* H = w.hits + (long long)p * kHitWords;
        if (threadIdx.x == 0) {
            int run = 0;
            for (int k = 0; k < kHitWords; ++k) {
                if (k == 32 && !vv) run = 0;  // EE block restarts at its own offset
                wpre[k] = run;
                if (!(vv && k >= 16)) run += __popcll(H[k]);
            }
        }
        __syncthreads();
        const int nw = vv ? 16 : kHitWords;
        for (int c = threadIdx.x; c < nw * 64; c += blockDim.x) {
            const int word = c >> 6, bit = c & 63;
            const unsigned long long m = H[word];
            if (!((m >> bit) & 1ull)) continue;
            const int rank = wpre[word] + __popcll(m & ((1ull << bit) - 1ull));
            const bool is_ee = !vv && c >= 2048;
            const int pos = (is_ee ? w.hit_off_ee[p] : w.hit_off_fv[p]) + rank;
            if (pos >= kMaxRawContacts) continue;
            if (vv) {
                const int kb = g.k[1];
                const int i = c / kb, j = c % kb;
                const V3 a = g.vA[i], b = g.vB[j];
                const V3 dd = b - a;
                write_contact(d, w, s, pos, g.A, g.B, 0, g.lvA[i], g.lvB[j], 0.5 * (a + b), dd, norm(dd), -2, -2);
            } else if (c < 1024) {
                const int kb = g.k[1];
                const int i = c / kb, j = c % kb;
                double sd;
                V3 proj;
                fv_test(g.fA[i], g.nA[i], g.vB[j], g.dhat, sd, proj);
                write_contact(d, w, s, pos, g.A, g.B, 0, g.lfA[i], g.lvB[j], 0.5 * (proj + g.vB[j]),
                              (sd >= 0.0) ? g.nA[i] : -g.nA[i], fabs(sd), -1, -1);
            } else if (c < 2048) {
                const int cc = c - 1024;
                const int ka = g.k[0];
                const int i = cc / ka, j = cc % ka;
                double sd;
                V3 proj;
                fv_test(g.fB[i], g.nB[i], g.vA[j], g.dhat, sd, proj);
                write_contact(d, w, s, pos, g.B, g.A, 0, g.lfB[i], g.lvA[j], 0.5 * (proj + g.vA[j]),
                              (sd >= 0.0) ? g.nB[i] : -g.nB[i], fabs(sd), -1, -1);
            } else {
                const int cc = c - 2048;
                const int kb = g.k[3];
                const int i = cc / kb, j = cc % kb;
                V3 c0, c1;
                closest_points(g.eA[i][0], g.eA[i][1], g.eB[j][0], g.eB[j][1], c0, c1);
                const V3 delta = c1 - c0;
                write_contact(d, w, s, pos, g.A, g.B, 1, g.leA[i][0], g.leB[j][0], 0.5 * (c0 + c1), delta,
                              norm(delta), g.leA[i][1], g.leB[j][1]);
            }
        }
        __syncthreads();
    }
}

// Truncation to 2000 contacts by distance (IPCSolver.cpp:571-577): std::sort of the raw
// list by distance, then the first 2000 are kept. With tied distances the kept set and
// its order are those of libstdc++'s introsort, so k_np_cap_order replays it on
// (distance, position) pairs by the wave, as the top-K lists do (tk_sort_blocks): the
// introsort loop's partitions, then its final insertion sort as per-block rank sorts,
// whose destinations below 2000 give kept[dst] = raw position. k_np_cap gathers them.
// (The per-scene scratch layout, kCap*B, and the gather, np_cap_gather, are shared with
// the ALM post-pass, which gathers in the ALM rounds.)

__global__ void __launch_bounds__(kTkLanes) k_np_cap_order(Dev d, Work w, int gate) {
    __shared__ int stk[3 * 64];
    const int s = blockIdx.x;
    const SceneParams& sp = d.sp[s];
    if (gate ? sp.alm_done : as_skip(w, sp)) return;
    const int n = w.raw_cnt[s];
    if (n <= kMaxContacts) return;
    const int lane = threadIdx.x;
    unsigned char* kb = reinterpret_cast<unsigned char*>(w.sort_key + (long long)s * kSceneScratch);
    unsigned char* ib = reinterpret_cast<unsigned char*>(w.sort_idx + (long long)s * kSceneScratch);
    TkHbm L;
    L.key = reinterpret_cast<double*>(kb);
    L.lpos = reinterpret_cast<unsigned short*>(kb + kCapLposB);
    L.rpos = reinterpret_cast<unsigned short*>(kb + kCapRposB);
    L.idx = reinterpret_cast<unsigned short*>(ib);
    L.start = ib + kCapStartB;
    L.stk = stk;
    int* kept = reinterpret_cast<int*>(ib + kCapKeptB);
    const long long cb = (long long)s * kMaxRawContacts;
    for (int i = lane; i < n; i += kTkLanes) {
        L.key[i] = w.cdbl[7 * (cb + i) + 6];
        L.idx[i] = (unsigned short)i;
    }
    __syncthreads();
    const KV16Store a{L.key, L.idx};
    tk_wave_sort_loop(a, n, L);
    // the final insertion sort: each element's rank within its block; the keys do not
    // move, so the destinations are computed in place
    for (int i = lane; i < n; i += kTkLanes) {
        int b0 = i, e = i + 1;
        while (!L.start[b0]) --b0;
        while (e < n && !L.start[e]) ++e;
        const double v = L.key[i];
        int r = b0;
        for (int t = b0; t < e; ++t) {
            const double u = L.key[t];
            r += (u < v) || (t < i && !(v < u));
        }
        if (r < kMaxContacts) kept[r] = L.idx[i];
    }
}

__global__ void __launch_bounds__(1024) k_np_cap(Dev d, Work w, int gate) {
    const int s = blockIdx.x;
    SceneParams& sp = d.sp[s];
    if (gate ? sp.alm_done : as_skip(w, sp)) return;
    np_cap_gather(w, sp, s);
}

void launch_narrow_phase(const Dev& d, const Work& w, int max_pairs, bool gate_alm, hipStream_t st, bool gather) {
    const int cap = w.grid_np > 0 ? w.grid_np : 4096;
    int grid = max_pairs < cap ? max_pairs : cap;
    if (grid < 1) grid = 1;
    const int gate = gate_alm ? 1 : 0;
    timing_begin("np_count", st);
    hipLaunchKernelGGL(k_np_count, dim3(grid), dim3(kBlock), 0, st, d, w, 0, gate);
    timing_end("np_count", st);
    hipLaunchKernelGGL(k_np_scan, dim3(d.S), dim3(kBlock), 0, st, d, w, 0, gate);
    hipLaunchKernelGGL(k_np_write, dim3(grid), dim3(kBlock), 0, st, d, w, 0, gate);
    // VV fallback for scenes without FV/EE contacts
    hipLaunchKernelGGL(k_np_count, dim3(grid), dim3(kBlock), 0, st, d, w, 1, gate);
    hipLaunchKernelGGL(k_np_scan, dim3(d.S), dim3(kBlock), 0, st, d, w, 1, gate);
    hipLaunchKernelGGL(k_np_write, dim3(grid), dim3(kBlock), 0, st, d, w, 1, gate);
    hipLaunchKernelGGL(k_np_cap_order, dim3(d.S), dim3(kTkLanes), 0, st, d, w, gate);
    if (gather) hipLaunchKernelGGL(k_np_cap, dim3(d.S), dim3(1024), 0, st, d, w, gate);
}

// ------------------------------------------------------------ scene min distance
// computeSceneMinDistance(bodies) < 1e-9 (AdaptiveBarrier.h:89-162): vertex of body i
// against every face of body j > i, plane distance with the barycentric test through
// the 3x3 cofactor inverse. A pair can only pass if its AABB gap is at most
// 1e-9 + 2e-6 * (longest face edge); pairs beyond the scene's prune margin are skipped.
// Within a pair, a face of J can only pass for a vertex of I if the face's box,
// inflated by 1e-8 + 1e-5 * (its longest edge), meets I's world AABB, and a vertex of I
// only if it lies in J's box inflated by the same bound for J's longest edge (the
// projection lies within the triangle enlarged by the 1e-6 barycentric tolerance).
// Both filters are conservative, so the flag is exact.
// The per-face quantities (unit normal, the inverse of [e1 e2 n]) are computed once per
// face, in chunks of kSmfChunk faces, and every (vertex, face) candidate of a pair is one
// work item: the same expressions as the per-vertex loop, evaluated once.
constexpr int kSmfChunk = 256;
// LDS of k_scene_min_flag, sized at launch by the largest mesh (NV vertices, NF faces):
// J's world vertices and I's candidate vertices (3 NV doubles each), the face chunk
// (v0, n, inverse: 15 doubles per face), its validity flags and J's active faces. For
// meshes above kTkLdsPrims (HBM = true) the vertex arrays and the active faces are in a
// per-work-group slab of HBM (smf_slab_bytes) and only the face chunk is in LDS.
__host__ __device__ inline size_t smf_lds_bytes(int NV, int NF) {
    return sizeof(double) * (6 * size_t(NV) + 15 * kSmfChunk) + sizeof(int) * (kSmfChunk + size_t(NF));
}
constexpr size_t kSmfChunkLds = sizeof(double) * 15 * kSmfChunk + sizeof(int) * kSmfChunk;
size_t smf_slab_bytes(int NV, int NF) {
    return (sizeof(double) * 6 * size_t(NV) + sizeof(int) * size_t(NF) + 255) & ~size_t(255);
}

template <bool HBM>
__global__ void __launch_bounds__(kBlock) k_scene_min_flag(Dev d, Work w, int ungated) {
    extern __shared__ __align__(16) unsigned char smf_lds[];
    const int NV = d.max_nv;
    double* wvx = reinterpret_cast<double*>(HBM ? w.smf_slab + size_t(blockIdx.x) * w.smf_slab_stride : smf_lds);
    double* wvy = wvx + NV;
    double* wvz = wvy + NV;
    double* ivx = wvz + NV;
    double* ivy = ivx + NV;
    double* ivz = ivy + NV;
    double* fd = HBM ? reinterpret_cast<double*>(smf_lds) : ivz + NV;
    int* fok = reinterpret_cast<int*>(fd + 15 * kSmfChunk);
    int* act = HBM ? reinterpret_cast<int*>(ivz + NV) : fok + kSmfChunk;
    __shared__ int nact, nvi;
    const int total = *w.n_prox_total;
    for (int p = blockIdx.x; p < total; p += gridDim.x) {
        const int s = w.prox_scene[p];
        const SceneParams& sp = d.sp[s];
        if (!ungated && !(sp.min_pair_dist < sp.dhat * 2.0)) continue;
        if (w.flag[s] || as_skip(w, sp)) continue;
        const int I = w.prox_pairs[2 * p], J = w.prox_pairs[2 * p + 1];
        const V3 pI = ld3(d.p, I), pJ = ld3(d.p, J);
        const Q qI = ld4(d.q, I), qJ = ld4(d.q, J);
        const int nvJ = d.nv[J];
        if (threadIdx.x == 0) {
            nact = 0;
            nvi = 0;
        }
        for (int k = threadIdx.x; k < nvJ; k += blockDim.x) {
            const V3 v = world_vertex(d, J, k, pJ, qJ);
            wvx[k] = v.x; wvy[k] = v.y; wvz[k] = v.z;
        }
        __syncthreads();
        const int nfJ = d.nf[J];
        const int* F = d.faces + 3 * d.f_off[J];
        const Box3 bI = box_inflated(ldbox(w.wbox, I), 1e-9);
        for (int f = threadIdx.x; f < nfJ; f += blockDim.x) {
            const int a = F[3 * f], b = F[3 * f + 1], c = F[3 * f + 2];
            const V3 A = v3(wvx[a], wvy[a], wvz[a]), B = v3(wvx[b], wvy[b], wvz[b]), C = v3(wvx[c], wvy[c], wvz[c]);
            const double L = dmax(norm(B - A), dmax(norm(C - B), norm(A - C)));
            Box3 fb;
            fb.mn = vmin(A, vmin(B, C));
            fb.mx = vmax(A, vmax(B, C));
            if (box_overlap(box_inflated(fb, 1e-8 + 1e-5 * L), bI)) act[atomicAdd(&nact, 1)] = f;
        }
        const Box3 bJ = box_inflated(ldbox(w.wbox, J), 1e-8 + 1e-5 * d.max_edge[J] + 1e-9);
        for (int vi = threadIdx.x; vi < d.nv[I]; vi += blockDim.x) {
            const V3 pt = world_vertex(d, I, vi, pI, qI);
            if (!(pt.x >= bJ.mn.x && pt.x <= bJ.mx.x && pt.y >= bJ.mn.y && pt.y <= bJ.mx.y && pt.z >= bJ.mn.z &&
                  pt.z <= bJ.mx.z))
                continue;
            const int k = atomicAdd(&nvi, 1);
            ivx[k] = pt.x; ivy[k] = pt.y; ivz[k] = pt.z;
        }
        __syncthreads();
        const int na = nact, nv = nvi;
        for (int c0 = 0; c0 < na && nv > 0; c0 += kSmfChunk) {
            const int nc = min(kSmfChunk, na - c0);
            for (int fi = threadIdx.x; fi < nc; fi += blockDim.x) {
                const int f = act[c0 + fi];
                const int a = F[3 * f], b = F[3 * f + 1], c = F[3 * f + 2];
                const V3 v0 = v3(wvx[a], wvy[a], wvz[a]);
                const V3 e1 = v3(wvx[b], wvy[b], wvz[b]) - v0;
                const V3 e2 = v3(wvx[c], wvy[c], wvz[c]) - v0;
                V3 n = cross(e1, e2);
                const double area = norm(n);
                fok[fi] = area < 1e-12 ? 0 : 1;
                n = normalized(n);
                M3 M;
                M.m[0] = e1.x; M.m[3] = e1.y; M.m[6] = e1.z;
                M.m[1] = e2.x; M.m[4] = e2.y; M.m[7] = e2.z;
                M.m[2] = n.x;  M.m[5] = n.y;  M.m[8] = n.z;
                const M3 Mi = m3inverse(M);
                double* o = fd + 15 * fi;
                o[0] = v0.x; o[1] = v0.y; o[2] = v0.z;
                o[3] = n.x;  o[4] = n.y;  o[5] = n.z;
                for (int k = 0; k < 9; ++k) o[6 + k] = Mi.m[k];
            }
            __syncthreads();
            // a wavefront takes a face (its data loaded once, a broadcast read) and its
            // lanes sweep the candidate vertices
            bool hit = false;
            const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
            for (int fi = wv; fi < nc; fi += kBlock / 64) {
                if (!fok[fi]) continue;
                const double* o = fd + 15 * fi;
                const V3 v0 = v3(o[0], o[1], o[2]), n = v3(o[3], o[4], o[5]);
                M3 Mi;
                for (int k = 0; k < 9; ++k) Mi.m[k] = o[6 + k];
                for (int vi = lane; vi < nv; vi += 64) {
                    const V3 r = v3(ivx[vi], ivy[vi], ivz[vi]) - v0;
                    const double dist = fabs(dot(r, n));
                    if (!(dist < 1e-9)) continue;
                    const V3 co = m3v(Mi, r);
                    const double b0 = 1.0 - co.x - co.y, b1 = co.x, b2 = co.y;
                    double bmn = b0, bmx = b0;
                    if (b1 < bmn) bmn = b1;
                    if (b2 < bmn) bmn = b2;
                    if (bmx < b1) bmx = b1;
                    if (bmx < b2) bmx = b2;
                    if (bmn >= -1e-6 && bmx <= 1.0 + 1e-6) hit = true;
                }
                if (__any(hit)) break;
            }
            if (__syncthreads_or(hit)) {
                if (threadIdx.x == 0) atomicOr(&w.flag[s], 1);
                break;
            }
        }
        __syncthreads();
    }
}

void launch_scene_min_flag(const Dev& d, const Work& w, int max_prox, hipStream_t st, int ungated) {
    int cap = w.grid_smf > 0 ? w.grid_smf : 2048;
    if (w.smf_slab && cap > w.smf_slab_blocks) cap = w.smf_slab_blocks;
    int grid = max_prox < cap ? max_prox : cap;
    if (grid < 1) grid = 1;
    if (w.smf_slab)
        hipLaunchKernelGGL(k_scene_min_flag<true>, dim3(grid), dim3(kBlock), kSmfChunkLds, st, d, w, ungated);
    else
        hipLaunchKernelGGL(k_scene_min_flag<false>, dim3(grid), dim3(kBlock), smf_lds_bytes(d.max_nv, d.max_nf), st,
                           d, w, ungated);
}

int geom_configure_smf() {
    // k_scene_min_flag<false> takes up to smf_lds_bytes(kTkLdsPrims, kTkLdsPrims) (~138 KiB)
    return hipFuncSetAttribute(reinterpret_cast<const void*>(k_scene_min_flag<false>),
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               int(smf_lds_bytes(kTkLdsPrims, kTkLdsPrims))) == hipSuccess
               ? 0
               : -1;
}

// ----------------------------------------------------------------------------- CCD
// Interval CCD (TimeOfImpact.cpp:8-150, IntervalRootFinder.h:20-114,
// RigidTrajectoryAABB.h:13-174): capped depth-first search (1000 pops), high half
// popped first, split of the first widest dimension at the round-to-nearest median,
// positions as the hull over the step, rotation by slerp at t.lo / t.hi.
struct CPose {
    V3 p0, p1;
    Q q0, q1;
};
struct IV3 {
    Iv x, y, z;
};

__device__ __forceinline__ IV3 vertex_traj(const CPose& ps, V3 lv, Iv t) {
    const Q qlo = slerp(ps.q0, t.lo, ps.q1), qhi = slerp(ps.q0, t.hi, ps.q1);
    const V3 a = rotate(qlo, lv), b = rotate(qhi, lv);
    IV3 r;
    r.x = iv(dmin(ps.p0.x, ps.p1.x), dmax(ps.p0.x, ps.p1.x)) + iv(dmin(a.x, b.x), dmax(a.x, b.x));
    r.y = iv(dmin(ps.p0.y, ps.p1.y), dmax(ps.p0.y, ps.p1.y)) + iv(dmin(a.y, b.y), dmax(a.y, b.y));
    r.z = iv(dmin(ps.p0.z, ps.p1.z), dmax(ps.p0.z, ps.p1.z)) + iv(dmin(a.z, b.z), dmax(a.z, b.z));
    return r;
}
__device__ __forceinline__ IV3 edge_traj(const CPose& ps, V3 a, V3 b, Iv t, Iv al) {
    const IV3 v0 = vertex_traj(ps, a, t), v1 = vertex_traj(ps, b, t);
    const Iv A = iv(al.lo, al.hi);
    IV3 r;
    r.x = v0.x * (1.0 - A) + v1.x * A;
    r.y = v0.y * (1.0 - A) + v1.y * A;
    r.z = v0.z * (1.0 - A) + v1.z * A;
    return r;
}
__device__ __forceinline__ IV3 face_traj(const CPose& ps, V3 a, V3 b, V3 c, Iv t, Iv u, Iv v) {
    const IV3 v0 = vertex_traj(ps, a, t), v1 = vertex_traj(ps, b, t), v2 = vertex_traj(ps, c, t);
    const Iv one = iv(1.0, 1.0);
    const Iv wi = (one - u) - v;
    IV3 r;
    r.x = (wi * v0.x + u * v1.x) + v * v2.x;
    r.y = (wi * v0.y + u * v1.y) + v * v2.y;
    r.z = (wi * v0.z + u * v1.z) + v * v2.z;
    return r;
}

// Query kinds 0-2 (edge-vertex, edge-edge, face-vertex) follow the reference. Kinds 3
// (face-vertex) and 4 (edge-edge) are the ipc mode's, written in body-relative form:
// the relative translation p_A(t) - p_B(t), linear in t over the box, plus each rotated
// body-frame point or edge vector as rot_hull, the slerp-endpoint hull widened by the
// arc's largest distance from its chord, |r| (1 - cos(phi / 2)) with cos(phi / 2) =
// |<q(t.lo), q(t.hi)>|, plus 8e-15 |r| for the rounding of the endpoint rotations.
__device__ __forceinline__ IV3 rot_hull(const CPose& ps, V3 r, Iv t) {
    const Q qlo = slerp(ps.q0, t.lo, ps.q1), qhi = slerp(ps.q0, t.hi, ps.q1);
    const V3 a = rotate(qlo, r), b = rotate(qhi, r);
    const double cd = dmin(1.0, fabs(((qlo.w * qhi.w + qlo.x * qhi.x) + qlo.y * qhi.y) + qlo.z * qhi.z));
    const double rl = sqrt((r.x * r.x + r.y * r.y) + r.z * r.z);
    const double e = rl * (1.0 - cd) + 8e-15 * rl;
    IV3 o;
    o.x = iv(sub_dn(dmin(a.x, b.x), e), add_up(dmax(a.x, b.x), e));
    o.y = iv(sub_dn(dmin(a.y, b.y), e), add_up(dmax(a.y, b.y), e));
    o.z = iv(sub_dn(dmin(a.z, b.z), e), add_up(dmax(a.z, b.z), e));
    return o;
}
__device__ __forceinline__ Iv rel_trans1(double a0, double a1, double b0, double b1, Iv t) {
    const Iv pa0 = iv(a0, a0), pb0 = iv(b0, b0);
    return (pa0 - pb0) + t * ((iv(a1, a1) - pa0) - (iv(b1, b1) - pb0));
}
__device__ __forceinline__ IV3 rel_trans(const CPose& A, const CPose& B, Iv t) {
    IV3 o;
    o.x = rel_trans1(A.p0.x, A.p1.x, B.p0.x, B.p1.x, t);
    o.y = rel_trans1(A.p0.y, A.p1.y, B.p0.y, B.p1.y, t);
    o.z = rel_trans1(A.p0.z, A.p1.z, B.p0.z, B.p1.z, t);
    return o;
}

struct CcdQuery {
    int kind;
    CPose A, B;
    V3 a[2], b[3];
};

__device__ __forceinline__ IV3 ccd_f(const CcdQuery& q, const Iv* x, bool& nan) {
    IV3 r;
    if (q.kind == 0) {
        const IV3 v = vertex_traj(q.A, q.a[0], x[0]);
        const IV3 e = edge_traj(q.B, q.b[0], q.b[1], x[0], x[1]);
        r.x = v.x - e.x; r.y = v.y - e.y; r.z = v.z - e.z;
    } else if (q.kind == 1) {
        const IV3 ea = edge_traj(q.A, q.a[0], q.a[1], x[0], x[1]);
        const IV3 eb = edge_traj(q.B, q.b[0], q.b[1], x[0], x[2]);
        r.x = ea.x - eb.x; r.y = ea.y - eb.y; r.z = ea.z - eb.z;
    } else if (q.kind == 2) {
        const IV3 v = vertex_traj(q.A, q.a[0], x[0]);
        const IV3 f = face_traj(q.B, q.b[0], q.b[1], q.b[2], x[0], x[1], x[2]);
        r.x = v.x - f.x; r.y = v.y - f.y; r.z = v.z - f.z;
    } else if (q.kind == 3) {
        // (p_A - p_B) + R_A a - R_B b0 - (u R_B (b1 - b0) + v R_B (b2 - b0))
        const Iv t = x[0], u = x[1], v = x[2];
        const IV3 tr = rel_trans(q.A, q.B, t), ra = rot_hull(q.A, q.a[0], t), rb = rot_hull(q.B, q.b[0], t);
        const IV3 r1 = rot_hull(q.B, q.b[1] - q.b[0], t), r2 = rot_hull(q.B, q.b[2] - q.b[0], t);
        r.x = ((tr.x + ra.x) - rb.x) - (u * r1.x + v * r2.x);
        r.y = ((tr.y + ra.y) - rb.y) - (u * r1.y + v * r2.y);
        r.z = ((tr.z + ra.z) - rb.z) - (u * r1.z + v * r2.z);
    } else {
        // (p_A - p_B) + R_A a0 + alpha R_A (a1 - a0) - (R_B b0 + beta R_B (b1 - b0))
        const Iv t = x[0], al = x[1], be = x[2];
        const IV3 tr = rel_trans(q.A, q.B, t), ra = rot_hull(q.A, q.a[0], t), rea = rot_hull(q.A, q.a[1] - q.a[0], t);
        const IV3 rb = rot_hull(q.B, q.b[0], t), reb = rot_hull(q.B, q.b[1] - q.b[0], t);
        r.x = ((tr.x + ra.x) + al * rea.x) - (rb.x + be * reb.x);
        r.y = ((tr.y + ra.y) + al * rea.y) - (rb.y + be * reb.y);
        r.z = ((tr.z + ra.z) + al * rea.z) - (rb.z + be * reb.z);
    }
    if (iv_nan(r.x) || iv_nan(r.y) || iv_nan(r.z)) nan = true;
    return r;
}

// Kind 2 keeps the reference's predicate (every barycentric weight certainly in [0, 1],
// which the root box u = v = [0, 1] fails). Kind 3 keeps a box that meets the triangle
// u, v >= 0, u + v <= 1: with u and v inside [0, 1] by construction, the box meets it
// when its lowest corner does, upper(1 - u.lo - v.lo) >= 0.
__device__ __forceinline__ bool ccd_valid(const CcdQuery& q, const Iv* x) {
    if (q.kind != 2 && q.kind != 3) return true;
    const Iv u = x[1], v = x[2];
    const Iv w = (iv(1.0, 1.0) - u) - v;
    if (q.kind == 3) return u.hi >= 0 && v.hi >= 0 && w.hi >= 0;
    return u.lo >= 0 && u.hi <= 1 && v.lo >= 0 && v.hi <= 1 && w.lo >= 0 && w.hi <= 1;
}

constexpr int kCcdStack = 100;

// Returns 1 on impact, 0 otherwise, -1 on NaN, -2 on stack overflow. The ipc mode's
// kinds 3 and 4 search earliest first (low half popped first), and a search stopped by
// the pop cap reports an impact at the least t.lo of its best leaf and of every box
// still on the stack (a lower bound of any impact it did not rule out).
__device__ int ccd_toi(const CcdQuery& q, double earliest, double tol, double* toi, int* pops) {
    const int D = q.kind == 0 ? 2 : 3;
    const bool early = q.kind >= 3;
    const double tl[3] = {tol, 1.0, 1.0};
    Iv stk[kCcdStack][3];
    int sp = 0;
    bool nan = false;
    Iv x0[3] = {iv(0.0, earliest), iv(0.0, 1.0), iv(0.0, 1.0)};
    {
        Iv xt[3] = {iv(0.0, tl[0]), iv(0.0, tl[1]), iv(0.0, tl[2])};
        (void)ccd_f(q, xt, nan);  // pre-check, result unused (IntervalRootFinder.h:41-48)
    }
    for (int k = 0; k < 3; ++k) stk[0][k] = x0[k];
    sp = 1;
    Iv best[3];
    best[0] = iv(1.0 / 0.0, 1.0 / 0.0);
    bool found = false;
    int it = 0;
    for (; sp > 0 && it < 1000; ++it) {
        Iv x[3];
        --sp;
        for (int k = 0; k < 3; ++k) x[k] = stk[sp][k];
        if (x[0].lo >= best[0].lo) continue;
        if (!ccd_valid(q, x)) continue;
        const IV3 y = ccd_f(q, x, nan);
        if (!(iv_zero_in(y.x) && iv_zero_in(y.y) && iv_zero_in(y.z))) continue;
        bool within = true;
        for (int k = 0; k < D; ++k)
            if (iv_width(x[k]) > tl[k]) { within = false; break; }
        if (within) {
            if (x[0].lo < best[0].lo) {
                for (int k = 0; k < 3; ++k) best[k] = x[k];
                found = true;
            }
            continue;
        }
        int sd = 0;
        double mw = iv_width(x[0]);
        for (int k = 1; k < D; ++k) {
            const double wk = iv_width(x[k]);
            if (wk > mw) { mw = wk; sd = k; }
        }
        if (sp + 2 > kCcdStack) { *pops = it; return -2; }
        const double mid = iv_median(x[sd]);
        for (int k = 0; k < 3; ++k) { stk[sp][k] = x[k]; stk[sp + 1][k] = x[k]; }
        // the half on top (sp + 1) is popped next: the high half, or for the ipc
        // mode's kinds the low half
        stk[sp][sd] = early ? iv(mid, x[sd].hi) : iv(x[sd].lo, mid);
        stk[sp + 1][sd] = early ? iv(x[sd].lo, mid) : iv(mid, x[sd].hi);
        sp += 2;
    }
    *pops = it;
    if (nan) return -1;
    if (early && it >= 1000 && sp > 0) {
        double m = best[0].lo;
        for (int k = 0; k < sp; ++k) m = dmin(m, stk[k][0].lo);
        best[0] = iv(m, m);
        found = true;
    }
    *toi = found ? best[0].lo : 1.0 / 0.0;
    return found ? 1 : 0;
}

// Outcome of ccd_toi for a face-vertex query. The root box has u = v = [0, 1], so the
// barycentric weight w = (1 - u) - v = [-1, 1] fails the domain check (lower bound
// exactly -1 under outward rounding) and the search ends at its first pop without an
// impact (TimeOfImpact.cpp:97-150 with IntervalRootFinder.h:20-114). What remains is
// the NaN screen of the pre-check evaluation, which decides between 0 and -1. Same
// result as the general search, without its explicit stack (no scratch traffic).
__device__ __forceinline__ int ccd_toi_face_vertex(const CcdQuery& q, double tol) {
    bool nan = false;
    const Iv xt[3] = {iv(0.0, tol), iv(0.0, 1.0), iv(0.0, 1.0)};
    (void)ccd_f(q, xt, nan);
    return nan ? -1 : 0;
}

__device__ CPose ld_cpose(const double* ps) {
    CPose p;
    p.p0 = v3(ps[0], ps[1], ps[2]);
    p.q0 = quat(ps[3], ps[4], ps[5], ps[6]);
    p.p1 = v3(ps[7], ps[8], ps[9]);
    p.q1 = quat(ps[10], ps[11], ps[12], ps[13]);
    return p;
}

__global__ void k_ccd_batch(int kind, int n, const double* pa, const double* pb, const double* a, const double* b,
                            const double* earliest, double tol, int* hit, double* toi, int* pops) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    CcdQuery q;
    q.kind = kind;
    q.A = ld_cpose(pa + 14 * i);
    q.B = ld_cpose(pb + 14 * i);
    const int na = (kind == 1 || kind == 4) ? 2 : 1, nb = (kind == 2 || kind == 3) ? 3 : 2;
    for (int k = 0; k < na; ++k) q.a[k] = ld3(a + 6 * i, k);
    for (int k = 0; k < nb; ++k) q.b[k] = ld3(b + 9 * i, k);
    double t = 1.0 / 0.0;
    int pp = 0;
    hit[i] = ccd_toi(q, earliest[i], tol, &t, &pp);
    toi[i] = t;
    pops[i] = pp;
}

void launch_ccd_batch(int kind, int n, const double* pa, const double* pb, const double* a, const double* b,
                      const double* earliest, double tol, int* hit, double* toi, int* pops, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_ccd_batch, dim3((n + 63) / 64), dim3(64), 0, st, kind, n, pa, pb, a, b, earliest, tol, hit,
                       toi, pops);
}

// CCDSystem::computeEarliestTOI (CCD.h:22-82) over the FV candidates of every close
// pair: vertex body moves p -> p + v dt with frozen rotation, face body likewise.
// Each candidate starts from earliest = 1; the scene result is the minimum. A
// face-vertex query never reports an impact (ccd_toi_face_vertex), so a step's outcome
// is only whether some query's pre-check evaluation meets a NaN.
//
// Screen (one thread per close pair): that evaluation is finite whenever every input
// is finite and at most 1e50 in magnitude (the two bodies' p, q, v, the step's dt and
// their mesh vertices). It is interval arithmetic (sums, and products with [-1, 1],
// [0, 1] or [0, 1e-8]) over the slerp of q with itself, the rotation q v q* (no
// division) and p + v dt: the slerp weights of a finite q are finite (linear weights
// when |q|^2 >= 1 - 2^-52, else sin ratios with sin(acos|q|^2) >= 2e-8), and every
// intermediate stays below ~1e160, far from overflow, so no inf or NaN can arise. Only
// pairs with an input outside that bound are listed for the evaluation itself; the
// outcome is that of evaluating every query.
// (w.ccd_bound is that bound, or negative to list every pair: NDIPC_CCD_EVAL_ALL=1 at
// upload.)
__device__ __forceinline__ bool ccd_inputs_bounded(const Dev& d, int b, double dt, double bound) {
    bool ok = d.vbound[b] <= bound && dt <= bound;
    for (int k = 0; k < 3; ++k) ok = ok && fabs(d.p[3 * b + k]) <= bound && fabs(d.v[3 * b + k]) <= bound;
    for (int k = 0; k < 4; ++k) ok = ok && fabs(d.q[4 * b + k]) <= bound;
    return ok;  // false for NaN
}

__global__ void k_ccd_reset(Work w) { *w.ccd_nlist = 0; }

__global__ void k_ccd_screen(Dev d, Work w, int gate_any) {
    const int total = *w.n_pairs_total;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < total; p += gridDim.x * blockDim.x) {
        const int* k = w.lcnt + p * 6;
        if (k[4] * k[1] + k[5] * k[0] == 0) continue;
        const SceneParams& sp = d.sp[w.pair_scene[p]];
        if (gate_any && !(sp.enable_ccd && sp.ccd_any)) continue;
        if (as_skip(w, sp)) continue;
        if (ccd_inputs_bounded(d, w.pairs[2 * p], sp.dt, w.ccd_bound) &&
            ccd_inputs_bounded(d, w.pairs[2 * p + 1], sp.dt, w.ccd_bound))
            continue;
        w.ccd_list[atomicAdd(w.ccd_nlist, 1)] = p;
    }
}

// The evaluation over the listed pairs. Work units of one wavefront: (listed pair, chunk
// of 64 of its at most 2048 face-vertex candidates), persistent over the units. A unit
// loads the pair's two step poses once and its lanes take the chunk's candidates (faces
// of A x vertices of B, then faces of B x vertices of A, from the top-K lists).
constexpr int kCcdChunks = 2 * kTopK * kTopK / 64;
__global__ void k_ccd_fv(Dev d, Work w) {
    const int total = *w.ccd_nlist;
    const int lane = threadIdx.x & 63;
    const long long wid = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nw = ((long long)gridDim.x * blockDim.x) >> 6;
    for (long long u = wid; u < (long long)total * kCcdChunks; u += nw) {
        const int p = w.ccd_list[int(u / kCcdChunks)], chunk = int(u % kCcdChunks);
        const int* k = w.lcnt + p * 6;
        const int k0 = k[0], k1 = k[1], n1 = k[4] * k1, n2 = k[5] * k0;
        const int local = chunk * 64 + lane;
        if (local >= n1 + n2) continue;
        SceneParams& sp = d.sp[w.pair_scene[p]];
        const int A = w.pairs[2 * p], B = w.pairs[2 * p + 1];
        const double dt = sp.dt;
        const bool first = local < n1;
        const int cc = first ? local : local - n1;
        const int fb = first ? A : B, vb = first ? B : A;
        const int fl = first ? w.lists[(p * 6 + 4) * kTopK + cc / k1] : w.lists[(p * 6 + 5) * kTopK + cc / k0];
        const int vl = first ? w.lists[(p * 6 + 1) * kTopK + cc % k1] : w.lists[(p * 6 + 0) * kTopK + cc % k0];
        CcdQuery q;
        q.kind = 2;
        q.A.p0 = ld3(d.p, vb); q.A.q0 = ld4(d.q, vb); q.A.p1 = q.A.p0 + ld3(d.v, vb) * dt; q.A.q1 = q.A.q0;
        q.B.p0 = ld3(d.p, fb); q.B.q0 = ld4(d.q, fb); q.B.p1 = q.B.p0 + ld3(d.v, fb) * dt; q.B.q1 = q.B.q0;
        q.a[0] = ld3(d.verts, d.v_off[vb] + vl);
        const int* f = d.faces + 3 * (d.f_off[fb] + fl);
        for (int m = 0; m < 3; ++m) q.b[m] = ld3(d.verts, d.v_off[fb] + f[m]);
        if (ccd_toi_face_vertex(q, 1e-8) < 0) atomicOr(&sp.solve_fail, 2);
    }
}

void launch_ccd_earliest(const Dev& d, const Work& w, int max_pairs, hipStream_t st) {
    timing_begin("ccd_fv", st);
    hipLaunchKernelGGL(k_ccd_reset, dim3(1), dim3(1), 0, st, w);
    const int gs = std::max(1, std::min((max_pairs + 255) / 256, 1024));
    hipLaunchKernelGGL(k_ccd_screen, dim3(gs), dim3(256), 0, st, d, w, 1);
    long long g = ((long long)max_pairs * kCcdChunks * 64 + 255) / 256;  // a wavefront per unit
    const long long cap = w.grid_ccd > 0 ? w.grid_ccd : 4096;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(k_ccd_fv, dim3((unsigned)g), dim3(256), 0, st, d, w);
    timing_end("ccd_fv", st);
}

__global__ void k_round_ops(const double* a, const double* b, int n, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[6 * i + 0] = add_dn(a[i], b[i]);
    out[6 * i + 1] = add_up(a[i], b[i]);
    out[6 * i + 2] = sub_dn(a[i], b[i]);
    out[6 * i + 3] = sub_up(a[i], b[i]);
    out[6 * i + 4] = mul_dn(a[i], b[i]);
    out[6 * i + 5] = mul_up(a[i], b[i]);
}

void launch_round_ops(const double* a, const double* b, int n, double* out, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_round_ops, dim3((n + 255) / 256), dim3(256), 0, st, a, b, n, out);
}

__global__ void k_elemfn(int fn, const double* x, const double* y, int n, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double r;
    switch (fn) {
        case 0: r = ef::log(x[i]); break;
        case 1: r = ef::sin(x[i]); break;
        case 2: r = ef::cos(x[i]); break;
        case 3: r = ef::atan2(x[i], y[i]); break;
        default: r = ef::acos(x[i]); break;
    }
    out[i] = r;
}

void launch_elemfn(int fn, const double* x, const double* y, int n, double* out, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_elemfn, dim3((n + 255) / 256), dim3(256), 0, st, fn, x, y, n, out);
}

}  // namespace nd
