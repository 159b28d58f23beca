// query.hip — closest-triangle queries on a resident LBVH.
//
// k_query_packet: wave-coherent traversal.  A wave (64 queries: a 4x4x4 brick of grid nodes, or 64
// consecutive points) walks the tree as one: the node index is wave-uniform, so each 64-byte node
// record arrives with one scalar load, every lane tests its own query against both child boxes,
// and two ballots decide whether to descend left, right or both (nearest-first by majority vote,
// the far child pushed on a wave-uniform stack held one entry per lane in a VGPR).  Popped entries
// are re-tested against each lane's current bound before the node is fetched again.  A leaf
// reference names a run of consecutive leaves (subtrees of at most kMaxRun leaves are never
// visited); the whole wave evaluates each triangle (point_triangle_dist_sq,
// serial/ssdf.hpp:50-121, in FP32), or skips it when its plane lies beyond every lane's bound.
// A query of more waves than the GPU holds at once runs as one resident set of blocks that take
// their waves from per-XCD queues (WaveSource).
//
// FIX (FP32 + FP64 fixup): pruning uses a widened bound, band(d) = (d (1 + tau) + e)^2 with e an
// FP32 error bound of a distance, and each lane lists the triangles inside that band.  After
// traversal, a lane whose band holds one triangle is final; two or more are re-evaluated in FP64
// (same formula, contraction off -> the oracle's bits) and the lowest index wins exact ties; a lane
// whose list overflowed is queued for the per-query kernel.
//
// k_coop: the queue's entries (those overflowed lists, and the queries of packet waves past their
// leaf budget), one wave per query: the exact FP64 search below the entry's FP32 band (plain mode:
// the FP32 search below its bound).
//
// k_query_exact: one lane per query, FP64 box and triangle tests, private stack; exact argmin with
// the canonical tie-break.  Used for FP64 meshes.
#include "query.hpp"
#include "radix.hpp"

#include <algorithm>
#include <atomic>
#include <climits>

namespace ssdf_amd {

constexpr int kWaves = kWavesPerBlock;
// FIX candidate list: 20 KB of LDS per block (5 KB per wave), the most that lets LDS allow 8 waves
// per SIMD (the packet kernels ask the compiler for 8: 64 VGPRs).  A query with more than kCand
// triangles in its band goes to the per-query kernel, which re-walks the tree in FP64.  Measured at
// 256^3 (FP32 + fixup): 200 k-triangle torus 7.9 / 7.8 / 8.4 ms with 16 / 20 / 24 entries; 2 M-
// triangle bumpy sphere 415 / 212 / 207 ms (1.5 M / 0.68 M / 0.32 M queries queued: inside a fine
// near-sphere, 17-24 triangles in a band are common).
constexpr int kCand = 20;
constexpr int kNone = INT_MIN;

__device__ __forceinline__ int uniform_i(int v) { return __builtin_amdgcn_readfirstlane(v); }

// v_writelane_b32: `old` with lane `lane` (wave-uniform) replaced by the wave-uniform `src`; one VALU
// operation where a lane compare and a select take two, plus the VGPR copy of src.  The ROCm 7.2
// clang has __builtin_amdgcn_readlane but no writelane builtin, so the declaration binds the LLVM
// intrinsic by name.
__device__ int amdgcn_writelane(int src, int lane, int old) __asm("llvm.amdgcn.writelane.i32");

// largest float below x (x > 0), so that "y <= prev_f(x)" is "y < x"
__device__ __forceinline__ float prev_f(float x) {
    if (x > 0.0f) return __int_as_float(__float_as_int(x) - 1);
    if (x == 0.0f) return -1.401298464e-45f;
    return x;
}

// (sqrt(d2) tau1 + babs)^2: the inclusive pruning bound around a squared distance.  The hardware
// square root (1 ulp, denormals may flush to 0) is enough: babs, 32 ulps of the coordinate scale,
// dwarfs its error.  +inf stays +inf.
__device__ __forceinline__ float band_r(float d2, float tau1, float babs) {
    return __builtin_amdgcn_sqrtf(d2) * tau1 + babs;
}
__device__ __forceinline__ float band_sq(float d2, float tau1, float babs) {
    const float r = band_r(d2, tau1, babs);
    return r * r;
}

// Sign of (q - c).n at the closest point c of triangle (a, b, c_); region by barycentrics with the
// reference's tolerance; follows serial/bvh/bvh.cpp:1734-1792.  Returns the dot product and |n|.
template <typename T>
__device__ __forceinline__ T sign_dot(const T qx, const T qy, const T qz, const T *A, const T *B, const T *C, int t,
                                      const int *__restrict__ s0, const int *__restrict__ s1,
                                      const int *__restrict__ s2, const double *__restrict__ face_n,
                                      const double *__restrict__ edge_n, const double *__restrict__ vertex_n,
                                      T &nlen) {
    T c[3];
    pt_tri_cp<T>(qx, qy, qz, A[0], A[1], A[2], B[0], B[1], B[2], C[0], C[1], C[2], c[0], c[1], c[2]);
    const T e1[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
    const T e2[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
    const T r[3] = {c[0] - A[0], c[1] - A[1], c[2] - A[2]};
    const T d11 = e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2];
    const T d12 = e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2];
    const T d22 = e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2];
    const T r1 = r[0] * e1[0] + r[1] * e1[1] + r[2] * e1[2];
    const T r2 = r[0] * e2[0] + r[1] * e2[1] + r[2] * e2[2];
    const T det = d11 * d22 - d12 * d12;
    const T l1 = det > T(0) ? (d22 * r1 - d12 * r2) / det : T(0);
    const T l2 = det > T(0) ? (d11 * r2 - d12 * r1) / det : T(0);
    const T l[3] = {T(1) - l1 - l2, l1, l2};
    const T tol = T(1e-4);
    int small = 0, last_small = -1, last_large = 0;
#pragma unroll
    for (int v = 0; v < 3; v++) {
        if (l[v] < tol)
            small++, last_small = v;
        else
            last_large = v;
    }
    const double *n;
    if (small == 0) {
        n = face_n + (size_t)3 * t;
    } else if (small == 1) {
        n = edge_n + (size_t)9 * t + 3 * ((last_small + 1) % 3);
    } else {
        const int vid = last_large == 0 ? s0[t] : (last_large == 1 ? s1[t] : s2[t]);
        n = vertex_n + (size_t)3 * vid;
    }
    const T n0 = (T)n[0], n1 = (T)n[1], n2 = (T)n[2];
    nlen = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
    return (qx - c[0]) * n0 + (qy - c[1]) * n1 + (qz - c[2]) * n2;
}


// Result of one FP32 query: leaf position k_final (< 0: nothing below the caller's bound), its FP32
// squared distance bestd, or its FP64 one d2_64 when have64; writes the output of mode OUT (SDF
// signs by the pseudonormal in FP32, re-evaluated in FP64 near the surface or when the FP32 dot
// product is too small to trust, so the sign is the FP64 one).  With have64 the output equals
// exact_write's for the same triangle.
template <int OUT>
__device__ __forceinline__ void packet_write(const QueryArgs &a, const long long qid, const float qx, const float qy,
                                             const float qz, const float out_init, const float init2,
                                             const float babs, const int k_final, const float bestd,
                                             const bool have64, const double d2_64) {
    const TriP *__restrict__ tris = a.tris;
    if (OUT == OUT_SQEDF) {
        if (k_final >= 0) {
            ((float *)a.out_d)[qid] = fminf(init2, have64 ? (float)d2_64 : bestd);
            if (a.out_tri) a.out_tri[qid] = tris[k_final].f.orig;
        }
        return;
    }
    const float dist = have64 ? (float)sqrt(d2_64) : sqrtf(bestd);
    if (OUT == OUT_EDF) {
        if (k_final >= 0) {
            ((float *)a.out_d)[qid] = fminf(out_init, dist);
            if (a.out_tri) a.out_tri[qid] = tris[k_final].f.orig;
        }
        return;
    }
    if (k_final < 0) {
        if (OUT == OUT_SDF) {
            ((float *)a.out_d)[qid] = SSDF_INF;
            if (a.out_tri) a.out_tri[qid] = -1;
        }
        return;
    }
    const TriF tr = tris[k_final].f;
    if (OUT == OUT_CLOSEST) {
        if (out_init > dist) {
            float cx, cy, cz;
            if (have64) {
                double ex, ey, ez;
                pt_tri_cp<double>((double)qx, (double)qy, (double)qz, tr.a[0], tr.a[1], tr.a[2], tr.b[0], tr.b[1],
                                  tr.b[2], tr.c[0], tr.c[1], tr.c[2], ex, ey, ez);
                cx = (float)ex, cy = (float)ey, cz = (float)ez;
            } else {
                pt_tri_cp<float>(qx, qy, qz, tr.a[0], tr.a[1], tr.a[2], tr.b[0], tr.b[1], tr.b[2], tr.c[0], tr.c[1],
                                 tr.c[2], cx, cy, cz);
            }
            ((float *)a.out_d)[qid] = dist;
            ((float *)a.out_cx)[qid] = cx;
            ((float *)a.out_cy)[qid] = cy;
            ((float *)a.out_cz)[qid] = cz;
            if (a.out_tri) a.out_tri[qid] = tr.orig;
        }
        return;
    }
    bool fp64 = (a.flags & QF_SIGN_FP64) || dist < a.sign_eps;
    bool neg = false;
    if (!fp64) {
        float nlen32;
        const float dot32 = sign_dot<float>(qx, qy, qz, tr.a, tr.b, tr.c, tr.orig, a.s0, a.s1, a.s2, a.face_n,
                                            a.edge_n, a.vertex_n, nlen32);
        neg = dot32 < 0.0f;
        fp64 = fabsf(dot32) <= (1e-3f * dist + 4.0f * babs) * nlen32;
    }
    if (fp64) {
        const double A[3] = {tr.a[0], tr.a[1], tr.a[2]}, B[3] = {tr.b[0], tr.b[1], tr.b[2]},
                     C[3] = {tr.c[0], tr.c[1], tr.c[2]};
        double nl64;
        const double dot64 = sign_dot<double>((double)qx, (double)qy, (double)qz, A, B, C, tr.orig, a.s0, a.s1,
                                              a.s2, a.face_n, a.edge_n, a.vertex_n, nl64);
        neg = dot64 < 0.0;
    }
    ((float *)a.out_d)[qid] = neg ? -dist : dist;
    if (a.out_tri) a.out_tri[qid] = tr.orig;
}

// Keeps the scalar loads of a whole triangle record ahead of the plane test: left alone, the
// compiler sinks the loads of the fields that only the full distance reads below the test's
// branch, so that a triangle that passes waits for a second round of scalar loads.  The pad word
// is named too, so that the record's 96 bytes after the header arrive in two loads, not three.
__device__ __forceinline__ void tri_loads_issued(const TriP &t) {
    asm volatile("" ::"s"(t.ux[0]), "s"(t.ux[1]), "s"(t.ux[2]), "s"(t.vx[0]), "s"(t.vx[1]), "s"(t.vx[2]));
    asm volatile("" ::"s"(t.bx), "s"(t.ibx), "s"(t.e1x), "s"(t.e1y), "s"(t.f1x), "s"(t.f1y));
    asm volatile("" ::"s"(t.cx), "s"(t.cy), "s"(t.g2x), "s"(t.g2y), "s"(t.in0), "s"(t.pad));
}

// One descent step's decision from the ballots of the node's two children (ml: lanes that want the
// left child, mr: the right one, mle: lanes nearer the left box), all in scalar code:
//   left first iff 2 |ml & (~mr | mle)| >= |ml | mr|: the majority of the wanting lanes is nearer to
//     it, and a child that is alone in the vote wins it (left alone: 2|ml| >= |ml|; right alone:
//     0 >= |mr| fails), so the taken child is wanted unless neither is;
//   ent  = node | side << 31, the child not taken (side 1: right);
//   sp  += 1 iff some lane wants the child not taken;
//   node = the child taken, or kNone when no lane wants either.
// Written out because the compiler keeps a uniform bool with several uses as a lane mask that each
// use re-tests (29 scalar instructions per visit instead of these 15).
__device__ __forceinline__ void descent_vote(const uint64_t ml, const uint64_t mr, const uint64_t mle, int &node,
                                             const int left, const int right, int &sp, int &ent) {
    uint64_t t, many;
    int a, b, next;
    asm("s_orn2_b64 %[t], %[mle], %[mr]\n\t"
        "s_and_b64 %[t], %[t], %[ml]\n\t"
        "s_bcnt1_i32_b64 %[a], %[t]\n\t"
        "s_or_b64 %[many], %[ml], %[mr]\n\t"
        "s_bcnt1_i32_b64 %[b], %[many]\n\t"
        "s_lshl_b32 %[a], %[a], 1\n\t"
        "s_cmp_ge_u32 %[a], %[b]\n\t"
        "s_cselect_b32 %[next], %[left], %[right]\n\t"
        "s_cselect_b64 %[t], %[mr], %[ml]\n\t"
        "s_cselect_b32 %[a], 0x80000000, 0\n\t"
        "s_or_b32 %[ent], %[node], %[a]\n\t"
        "s_cmp_lg_u64 %[t], 0\n\t"
        "s_addc_u32 %[sp], %[sp], 0\n\t"
        "s_cmp_lg_u64 %[many], 0\n\t"
        "s_cselect_b32 %[node], %[next], 0x80000000"
        : [t] "=&s"(t), [many] "=&s"(many), [a] "=&s"(a), [b] "=&s"(b), [ent] "=&s"(ent), [next] "=&s"(next),
          [sp] "+s"(sp), [node] "+s"(node)
        : [ml] "s"(ml), [mr] "s"(mr), [mle] "s"(mle), [left] "s"(left), [right] "s"(right)
        : "scc");
}

// descent_vote with the push folded in, for one-VGPR stacks: the other child's entry goes to lane
// sp (the depth before the vote) of stackv by v_writelane, lane select m0 (a VOP3 reads one SGPR, and
// the entry is in one), so that the depth is updated in place instead of through copies.
__device__ __forceinline__ void descent_vote_push(const uint64_t ml, const uint64_t mr, const uint64_t mle, int &node,
                                                  const int left, const int right, int &sp, int &stackv) {
    uint64_t t, many;
    int a, b, next, ent;
    asm("s_mov_b32 m0, %[sp]\n\t"
        "s_orn2_b64 %[t], %[mle], %[mr]\n\t"
        "s_and_b64 %[t], %[t], %[ml]\n\t"
        "s_bcnt1_i32_b64 %[a], %[t]\n\t"
        "s_or_b64 %[many], %[ml], %[mr]\n\t"
        "s_bcnt1_i32_b64 %[b], %[many]\n\t"
        "s_lshl_b32 %[a], %[a], 1\n\t"
        "s_cmp_ge_u32 %[a], %[b]\n\t"
        "s_cselect_b32 %[next], %[left], %[right]\n\t"
        "s_cselect_b64 %[t], %[mr], %[ml]\n\t"
        "s_cselect_b32 %[a], 0x80000000, 0\n\t"
        "s_or_b32 %[ent], %[node], %[a]\n\t"
        "s_cmp_lg_u64 %[t], 0\n\t"
        "s_addc_u32 %[sp], %[sp], 0\n\t"
        "s_cmp_lg_u64 %[many], 0\n\t"
        "s_cselect_b32 %[node], %[next], 0x80000000\n\t"
        "v_writelane_b32 %[stk], %[ent], m0"
        : [t] "=&s"(t), [many] "=&s"(many), [a] "=&s"(a), [b] "=&s"(b), [ent] "=&s"(ent), [next] "=&s"(next),
          [sp] "+s"(sp), [node] "+s"(node), [stk] "+v"(stackv)
        : [ml] "s"(ml), [mr] "s"(mr), [mle] "s"(mle), [left] "s"(left), [right] "s"(right)
        : "scc", "m0");
}

// Byte b of the node array.  STACK == 64 launches serve trees of fewer than kNarrowNodes internal
// nodes (the launch picks STACK from the depth and the size), whose byte offsets fit 32 bits: the
// scalar loads then take the offset as their SGPR offset (one shift instead of a 64-bit multiply-add);
// STACK == 128 launches index in 64 bits.
template <int STACK, typename T>
__device__ __forceinline__ const T &node_bytes(const Node *__restrict__ nodes, const unsigned long long b) {
    return STACK == 64 ? *(const T *)((const char *)nodes + (unsigned)b) : *(const T *)((const char *)nodes + b);
}

// Stack entry i (parent | side << 31): the pending child's box and reference, read from the parent
// record with scalar loads issued together (the reference is not left for after the box test).
template <int STACK>
__device__ __forceinline__ void stack_entry_box(const Node *__restrict__ nodes, const int stackv, const int stackv_hi,
                                                const int i, float *mn, float *mx, int &child) {
    const int e = (STACK == 64 || i < 64) ? __builtin_amdgcn_readlane(stackv, i & 63)
                                          : __builtin_amdgcn_readlane(stackv_hi, i & 63);
    const unsigned side = (unsigned)e >> 31;
    // the record's byte offset: 32-bit e << 6 drops the side bit (STACK == 64 trees have fewer than
    // 2^26 internal nodes), 64-bit otherwise
    const unsigned long long rec =
        STACK == 64 ? (unsigned long long)((unsigned)e << 6) : (unsigned long long)((unsigned)e & 0x7fffffffu) << 6;
    struct Box6 {
        float v[6];
    };
    const Box6 bx = node_bytes<STACK, Box6>(nodes, STACK == 64 ? (unsigned)rec + (side ? 24u : 0u) : rec + 24 * side);
#pragma unroll
    for (int d = 0; d < 3; d++) mn[d] = bx.v[d], mx[d] = bx.v[3 + d];
    child = uniform_i(node_bytes<STACK, int>(nodes, STACK == 64 ? (unsigned)rec + (side ? 52u : 48u) : rec + 48 + 4 * side));
}

// ------------------------------------------------------------------------------------------------
// Wave-coherent FP32 traversal (float queries, float meshes).
//
// Stack: wave-uniform ints, entry i in lane i % 64 of a VGPR, each naming where the pending subtree
// hangs (parent | side << 31); a pop re-reads that child's box and reference from the parent record
// with scalar loads and re-tests the box.
// FIX candidates: per lane, an append-only list of leaf positions in LDS ([kCand][256], lane-major,
// conflict-free).  A triangle enters when its FP32 d2 is inside band(best); a new best whose own
// band lies below the old best empties the list first (every old entry is >= the old best, so
// none of them can be in the new band), and a full list drops the entries outside the current
// band.  The list therefore always holds every triangle inside the current band, plus possibly a
// few outside it, which the FP64 evaluation rejects.  A list that stays full after the drop (more
// than kCand triangles inside the band) sends the query to the exact FP64 traversal.
// ------------------------------------------------------------------------------------------------
// One wave's 64 queries (wave_id of the query source); `cand` is the block's LDS candidate table.
template <int STACK, int QSRC, int OUT, bool FIX>
__device__ __forceinline__ void packet_wave(const Node *__restrict__ nodes, const TriP *__restrict__ tris,
                                            const QueryArgs &a, const unsigned wave_id, int (*cand)[256]) {
    static_assert(STACK == 64 || STACK == 128, "the traversal stack lives in one or two VGPRs");
    const int lane = threadIdx.x & 63;
    long long qid = 0;
    float qx = 0.f, qy = 0.f, qz = 0.f;
    const bool valid = lane_query<QSRC, float>(a, wave_id, lane, qid, qx, qy, qz);
    if (__ballot(valid) == 0) return;
    bool active = valid;  // false once handed to the cooperative kernel

    // FP32 error bound of a distance: 2^-18 (32 ulps of 1) times the largest coordinate involved
    const float tau1 = 1.0f + a.tau;
    const float babs = fmaxf(a.band_abs, 0x1p-18f * fmaxf(fabsf(qx), fmaxf(fabsf(qy), fabsf(qz))));

    // The caller's current distance bounds the search strictly (EDF / CLOSEST / SQEDF).  Plain mode
    // keeps `best` and prunes with box < best (the first triangle found wins an exact FP32 tie);
    // FIX keeps an inclusive bound thr >= best, its near-tie band.
    float out_init = SSDF_INF;
    float init2 = SSDF_INF;
    if ((OUT == OUT_EDF || OUT == OUT_CLOSEST) && valid) {
        out_init = ((const float *)a.out_d)[qid];
        init2 = out_init * out_init;
    }
    if (OUT == OUT_SQEDF && valid) init2 = ((const float *)a.out_d)[qid];  // the seed is already squared
    float best = valid ? init2 : -SSDF_INF;  // plain: strict bound (invalid lanes never pass)
    const float init_thr = __builtin_bit_cast(float, __builtin_bit_cast(int, init2) - 1);  // largest below init2
    float thr = valid ? init_thr : -SSDF_INF;  // FIX: inclusive bound
    int bk = -1, ncand = 0;
    float bestd = SSDF_INF;  // FIX: FP32 best distance found
    // Plane bound of the lane's pruning bound (plane_bound): a triangle whose plane lies farther than
    // this from every lane's query is skipped by the whole wave after 7 of its ~40 instructions.
    float rb = valid ? plane_bound(__builtin_amdgcn_sqrtf(init2), babs) : -SSDF_INF;


    // Stack of pending subtrees: entry i sits in lane i % 64 of VGPR i / 64 (written by v_writelane
    // at the wave-uniform depth sp, read back with v_readlane), each naming where the subtree hangs:
    // parent | side << 31.
    int stackv = 0, stackv_hi = 0;
    int sp = 0;
    bool overflow = false;
    int iters = 0;
    const int max_iters = (int)a.max_iters;
    int node = uniform_i(a.meta->root);
    // The leaf budget (leaf references, a run counting once) shares the loop's one iteration test
    // (lim): the hand-over below stays out of the loop body, where it cost the loop 5-10% even when
    // never taken.
    int lim = a.heavy_iters > 0 ? min(max_iters, a.heavy_iters) : max_iters;
    for (;;) {
    for (;;) {
        // descend through internal nodes, nearest child first by majority vote; the decision and the
        // push are uniform selects, so a visit is straight-line code
        while (node >= 0) {
            const Node nd = node_bytes<STACK, Node>(nodes, (unsigned long long)node << 6);
            // the whole 64-byte record in one scalar load (first and last are not read by the walk)
            asm volatile("" ::"s"(nd.first), "s"(nd.last));
            const float dl = box_d2_fast(qx, qy, qz, nd.lmin, nd.lmax);
            const float dr = box_d2_fast(qx, qy, qz, nd.rmin, nd.rmax);
            const uint64_t ml = __ballot(FIX ? dl <= thr : dl < best);
            const uint64_t mr = __ballot(FIX ? dr <= thr : dr < best);
            const uint64_t mle = __ballot(dl <= dr);
            // the other child's entry is written at the old depth whether pushed or not (above the
            // top it is overwritten before any pop reads it), by v_writelane from a scalar.  The
            // depth is at most this node's depth < STACK: the stack holds at most one entry per
            // level above the node, and the launch picks STACK from the tree depth.
            if (STACK == 64) {
                descent_vote_push(ml, mr, mle, node, nd.left, nd.right, sp, stackv);
            } else {
                const int sp0 = sp;
                int ent;
                descent_vote(ml, mr, mle, node, nd.left, nd.right, sp, ent);
                if (sp0 < 64)
                    stackv = amdgcn_writelane(ent, sp0 & 63, stackv);
                else
                    stackv_hi = amdgcn_writelane(ent, sp0 & 63, stackv_hi);
            }
        }
        overflow |= sp > STACK;
        // The pop that follows re-reads the top entry's box and child reference; those scalar loads
        // are issued here, before the leaf's, so that the two latencies overlap.  The box is tested
        // after the leaf, against the bound the leaf may have tightened.
        float pmn[3] = {0.f, 0.f, 0.f}, pmx[3] = {0.f, 0.f, 0.f};
        int pchild = kNone;
        if (sp > 0) stack_entry_box<STACK>(nodes, stackv, stackv_hi, sp - 1, pmn, pmx, pchild);
        if (node != kNone) {
          // a run of leaves (one, or the leaves of a subtree too small to visit)
          const int k0 = leaf_first(node), nlv = leaf_count(node);
          const TriP *__restrict__ run = tris + k0;  // the run's records by 32-bit offsets from here
          for (int j = 0; j < nlv; j++) {
            const int k = k0 + j;
            const TriP tr = *(const TriP *)((const char *)run + (unsigned)j * (unsigned)sizeof(TriP));
            const TriPlane pl = tri_plane(qx, qy, qz, tr);
            tri_loads_issued(tr);
            if (__ballot(fabsf(pl.h) <= rb) == 0) continue;
            const float d2 = tri_d2_rest(qx, qy, qz, tr, pl);
            if (!FIX) {
                if (d2 < best) best = d2, bk = k, rb = plane_bound(__builtin_amdgcn_sqrtf(d2), babs);
            } else if (d2 <= thr) {
                if (d2 < bestd) {
                    const float r = band_r(d2, tau1, babs), nb = r * r;
                    if (nb < bestd) ncand = 0;
                    bestd = d2;
                    bk = k;
                    thr = fminf(nb, init_thr);
                    rb = plane_bound(r, babs);  // r^2 rounds to nb >= thr
                }
                if (ncand < kCand) {
                    cand[ncand][threadIdx.x] = k;
                    ncand++;
                } else if (ncand == kCand) {
                    // full: drop the entries the band has shrunk past (re-evaluated by the same FP32
                    // routine on the same inputs, so with the same values they entered with)
                    int keep = 0;
                    for (int i = 0; i < kCand; i++) {
                        const int kc = cand[i][threadIdx.x];
                        if (tri_d2_fast(qx, qy, qz, tris[kc]) <= thr) cand[keep++][threadIdx.x] = kc;
                    }
                    if (keep < kCand) cand[keep][threadIdx.x] = k;
                    ncand = keep + 1;  // kCand + 1: more than kCand inside the band, exact path
                }
            }
          }
        }
        // pop the nearest pending subtree some lane still needs (entry sp - 1 is in pmn/pmx/pchild)
        node = kNone;
        while (sp > 0) {
            --sp;
            const float dd = box_d2_fast(qx, qy, qz, pmn, pmx);
            if (__ballot(FIX ? dd <= thr : dd < best) != 0) {
                node = pchild;
                break;
            }
            if (sp > 0) stack_entry_box<STACK>(nodes, stackv, stackv_hi, sp - 1, pmn, pmx, pchild);
        }
        if (node == kNone || overflow) break;
        if (++iters > lim) break;
    }
    if (node == kNone || overflow || iters > max_iters) break;
    lim = max_iters;
    {
        // Past the leaf budget the wave's queries disagree too much for one shared walk: each
        // still active query goes to the queue with the bound found so far (FIX: the FP32 band of
        // its best; plain: its best) and is searched by a wave of its own.  Queries that find the
        // hand-over share of the queue full stay here, and the walk goes on.
        const uint64_t am = __ballot(active);
        unsigned base = 0, got = 0;
        if (lane == 0) {
            // reserve up to the hand-over share of the queue, never past it
            const unsigned want = (unsigned)__popcll(am);
            unsigned cur = __hip_atomic_load(a.fix_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            while (cur < a.heavy_cap) {
                const unsigned take = min(want, a.heavy_cap - cur);
                const unsigned prev = atomicCAS(a.fix_count, cur, cur + take);
                if (prev == cur) {
                    base = cur, got = take;
                    break;
                }
                cur = prev;
            }
        }
        base = (unsigned)__shfl((int)base, 0);
        got = (unsigned)__shfl((int)got, 0);
        const unsigned rank = (unsigned)__popcll(am & ((1ull << lane) - 1ull));
        if (active && rank < got) {
            const unsigned slot = base + rank;
            a.fix_list[4 * slot] = (unsigned)qid;
            // the bound handed over is the FP32 band of the best so far (plain mode too: the
            // per-query walk then finds that triangle again even where an ancestor box's FP32
            // distance rounds above the triangle's)
            // (FIX with a sliver flagged degenerate as the best so far: no bound, since the FP64
            // routine's value for such a triangle can exceed its FP32 one by more than the band)
            const bool sliver = FIX && bk >= 0 && (tris[bk].f.flags & 1);
            a.fix_list[4 * slot + 1] = __float_as_uint(sliver ? SSDF_INF : band_sq(FIX ? bestd : best, tau1, babs));
            a.fix_list[4 * slot + 2] = a.fix_tag;
            a.fix_list[4 * slot + 3] = (unsigned)((unsigned long long)qid >> 32);
            active = false;
            best = -SSDF_INF;
            thr = -SSDF_INF;
            rb = -SSDF_INF;
        }
        if (__ballot(active) == 0) return;
    }
    }
    if (lane == 0 && (overflow || iters > max_iters)) atomicOr(a.err_flag, overflow ? 1 : 2);
    if (!FIX) bestd = best;
    if (!active) return;

    // ---- resolve the closest triangle ----
    int k_final = bk;      // leaf position
    bool have64 = false;   // distance known in FP64
    double d2_64 = 0.0;
    bool unbounded = false;  // re-search in FP64 without the FP32 band
    if (FIX && bk >= 0) {
        bool exact_needed = ncand > kCand;
        if (!exact_needed && (ncand >= 2 || (tris[bk].f.flags & 1))) {
            const double px = qx, py = qy, pz = qz;
            // On slivers flagged degenerate (height below 1e-6 of the longest edge) the FP64
            // routine's value can exceed the true distance, which the FP32 one follows, by more than
            // the band: the list then need not hold the FP64 argmin, so such a query is searched
            // again in FP64 without a bound.
            const double band64 = (double)band_sq(bestd, tau1, babs);
            double bestd = __builtin_inf();
            int bestt = INT_MAX, bestk = -1;
            for (int i = 0; i < ncand; i++) {
                const int kc = cand[i][threadIdx.x];
                const TriF tr = tris[kc].f;
                const double d = pt_tri_d2<double>(px, py, pz, tr.a[0], tr.a[1], tr.a[2], tr.b[0], tr.b[1], tr.b[2],
                                                   tr.c[0], tr.c[1], tr.c[2]);
                if ((tr.flags & 1) && d > band64) unbounded = true;
                if (d < bestd || (d == bestd && tr.orig < bestt)) bestd = d, bestt = tr.orig, bestk = kc;
            }
            if (unbounded) {
                exact_needed = true;
            } else if (bestk >= 0) {
                k_final = bestk;
                have64 = true;
                d2_64 = bestd;
            } else {
                exact_needed = true;  // FP64 rejects every candidate (degenerate triangles)
            }
        } else if (!exact_needed) {
            // one candidate: its FP64 distance, so that the output is a function of the query and
            // the canonical triangle alone, whichever path found it
            const TriF tr = tris[bk].f;
            d2_64 = pt_tri_d2<double>((double)qx, (double)qy, (double)qz, tr.a[0], tr.a[1], tr.a[2], tr.b[0],
                                      tr.b[1], tr.b[2], tr.c[0], tr.c[1], tr.c[2]);
            have64 = true;
        }
        if (exact_needed) {
            // the exact kernel recomputes this query from the untouched output (its seed), searching
            // only the FP32 band: the FP64 minimum cannot lie outside it
            const unsigned pos = atomicAdd(a.fix_count, 1u);
            if (pos < a.fix_cap) {
                a.fix_list[4 * pos] = (unsigned)qid;
                a.fix_list[4 * pos + 1] = __float_as_uint(unbounded ? SSDF_INF : band_sq(bestd, tau1, babs));
                a.fix_list[4 * pos + 2] = a.fix_tag;
                a.fix_list[4 * pos + 3] = (unsigned)((unsigned long long)qid >> 32);
                return;
            }
            // queue full: keep the FP32 result and report it (error bit 4)
            atomicOr(a.err_flag, 4);
        }
    }

    packet_write<OUT>(a, qid, qx, qy, qz, out_init, init2, babs, k_final, bestd, have64, d2_64);
}

template <int STACK, int QSRC, int OUT, bool FIX>
// 8 waves per SIMD (64 VGPRs): FIX spills about 60 VGPRs (the grid and point kernels), stored and
// reloaded once per wave around the FP64 candidate evaluation and the hand-over, none in the
// traversal loops (descent, leaf, pop).  Before the persistent launch it spilled 17 and still
// gained 8% on the 200 k-triangle torus at 256^3 over its unconstrained 78 VGPRs and 6 waves.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_query_packet(const Node *__restrict__ nodes, const TriP *__restrict__ tris,
                                                      QueryArgs a) {
    __shared__ int cand[FIX ? kCand : 1][256];
    // one call site of the walk for both launch kinds (inlined twice, it spilled three times as much)
    WaveSource src(a);
    for (unsigned w; src.next(a, w);) packet_wave<STACK, QSRC, OUT, FIX>(nodes, tris, a, w, cand);
}

// ------------------------------------------------------------------------------------------------
// Exact FP64 traversal, one lane per query.  A box is skipped only when its FP64 distance exceeds
// the best distance by delta64 (far above FP64 rounding of any distance), so the result is the
// brute-force argmin of the FP64 squared distance with the lowest index on exact ties.
// ------------------------------------------------------------------------------------------------
template <typename G>
__device__ __forceinline__ void tri_verts64(const QueryArgs &a, const TriF &tr, double *A, double *B, double *C) {
    if (sizeof(G) == sizeof(float)) {
#pragma unroll
        for (int d = 0; d < 3; d++) A[d] = tr.a[d], B[d] = tr.b[d], C[d] = tr.c[d];
    } else {
        const int t = tr.orig;
        const int i0 = a.s0[t], i1 = a.s1[t], i2 = a.s2[t];
        const double *sx = (const double *)a.sx, *sy = (const double *)a.sy, *sz = (const double *)a.sz;
        A[0] = sx[i0], A[1] = sy[i0], A[2] = sz[i0];
        B[0] = sx[i1], B[1] = sy[i1], B[2] = sz[i1];
        C[0] = sx[i2], C[1] = sy[i2], C[2] = sz[i2];
    }
}

constexpr int kExactStack = 128;

template <typename G, int OUT>
__device__ void exact_write(const QueryArgs &a, long long qid, double qx, double qy, double qz, double out_init,
                            double best, int bt, int bk);

// One query, exact FP64: qid of the query source; band2 bounds the search (the FP32 band of a
// queued fixup query, +inf otherwise).
// Coordinates of query qid in caller precision G, widened to FP64.
template <typename G, int QSRC>
__device__ __forceinline__ void query_coords(const QueryArgs &a, const long long qid, double &qx, double &qy,
                                             double &qz) {
    if (QSRC == QSRC_POINTS) {
        qx = (double)load_as<G>(a.qx, qid), qy = (double)load_as<G>(a.qy, qid), qz = (double)load_as<G>(a.qz, qid);
    } else {
        const long long i = qid % a.nx, j = (qid / a.nx) % a.ny, k = qid / ((long long)a.nx * a.ny);
        const G hx = (G)a.gh[0], hy = (G)a.gh[1], hz = (G)a.gh[2];
        qx = (double)((G)a.gmin[0] + (G)i * hx);
        qy = (double)((G)a.gmin[1] + (G)j * hy);
        qz = (double)((G)a.gmin[2] + (G)k * hz);
    }
}

template <typename G, int QSRC, int OUT>
__device__ void exact_one(const QueryArgs &a, const long long qid, const double band2) {
    double qx, qy, qz;
    query_coords<G, QSRC>(a, qid, qx, qy, qz);
    double out_init = __builtin_inf();
    if (OUT == OUT_EDF || OUT == OUT_CLOSEST || OUT == OUT_SQEDF) out_init = (double)load_as<G>(a.out_d, qid);
    const double init2 = OUT == OUT_SQEDF ? out_init : out_init * out_init;  // SQEDF seeds are squared
    double best = __builtin_inf(), bound = __builtin_inf();
    int bt = INT_MAX, bk = -1;
    const double delta = fmax(a.delta64, 1e-13 * fmax(fabs(qx), fmax(fabs(qy), fabs(qz))));
    {
        const double b0 = sqrt(fmin(init2, band2)) + delta;
        bound = b0 * b0;
    }
    int st_node[kExactStack];
    double st_d[kExactStack];
    int sp = 0;
    int node = a.meta->root;
    long long iters = 0;
    for (;;) {
        if (++iters > a.max_iters) {
            atomicOr(a.err_flag, 2);
            break;
        }
        if (node >= 0) {
            const Node nd = a.nodes[node];
            const double dl = pt_box_d2<double>(qx, qy, qz, nd.lmin[0], nd.lmin[1], nd.lmin[2], nd.lmax[0],
                                                nd.lmax[1], nd.lmax[2]);
            const double dr = pt_box_d2<double>(qx, qy, qz, nd.rmin[0], nd.rmin[1], nd.rmin[2], nd.rmax[0],
                                                nd.rmax[1], nd.rmax[2]);
            const bool nl = dl <= bound, nr = dr <= bound;
            if (nl && nr) {
                const bool lfirst = dl <= dr;
                if (sp < kExactStack) {
                    st_node[sp] = lfirst ? nd.right : nd.left;
                    st_d[sp] = lfirst ? dr : dl;
                    sp++;
                } else {
                    atomicOr(a.err_flag, 1);
                }
                node = lfirst ? nd.left : nd.right;
                continue;
            }
            if (nl) {
                node = nd.left;
                continue;
            }
            if (nr) {
                node = nd.right;
                continue;
            }
        } else {
            const int k0 = leaf_first(node), nlv = leaf_count(node);  // a run of leaves
            for (int j = 0; j < nlv; j++) {
                const TriF tr = a.tris[k0 + j].f;
                double A[3], B[3], C[3];
                tri_verts64<G>(a, tr, A, B, C);
                const double d =
                    pt_tri_d2<double>(qx, qy, qz, A[0], A[1], A[2], B[0], B[1], B[2], C[0], C[1], C[2]);
                if (d < init2 && (d < best || (d == best && tr.orig < bt))) {
                    best = d, bt = tr.orig, bk = k0 + j;
                    const double r = sqrt(best) + delta;
                    bound = r * r;
                }
            }
        }
        node = kNone;
        while (sp > 0) {
            --sp;
            if (st_d[sp] <= bound) {
                node = st_node[sp];
                break;
            }
        }
        if (node == kNone) break;
    }
    exact_write<G, OUT>(a, qid, qx, qy, qz, out_init, best, bt, bk);
}

// Writes the result of an exact search (FP64 best d2, original index bt, leaf position bk; bk < 0:
// nothing below the caller's bound) in caller precision G.
template <typename G, int OUT>
__device__ void exact_write(const QueryArgs &a, const long long qid, const double qx, const double qy,
                            const double qz, const double out_init, const double best, const int bt, const int bk) {
    if (OUT == OUT_SQEDF) {  // out_init holds the squared seed
        if (bk >= 0) {
            store_as<G>(a.out_d, qid, (G)fmin(out_init, best));
            if (a.out_tri) a.out_tri[qid] = bt;
        }
        return;
    }
    const double dist = sqrt(best);
    if (OUT == OUT_EDF) {
        if (bk >= 0) {
            store_as<G>(a.out_d, qid, (G)fmin(out_init, dist));
            if (a.out_tri) a.out_tri[qid] = bt;
        }
        return;
    }
    if (bk < 0) {
        if (OUT == OUT_SDF) {
            store_as<G>(a.out_d, qid, (G)__builtin_inf());
            if (a.out_tri) a.out_tri[qid] = -1;
        }
        return;
    }
    const TriF tr = a.tris[bk].f;
    double A[3], B[3], C[3];
    tri_verts64<G>(a, tr, A, B, C);
    if (OUT == OUT_CLOSEST) {
        const G gd = (G)dist;
        if ((G)out_init > gd) {
            double cx, cy, cz;
            pt_tri_cp<double>(qx, qy, qz, A[0], A[1], A[2], B[0], B[1], B[2], C[0], C[1], C[2], cx, cy, cz);
            store_as<G>(a.out_d, qid, gd);
            store_as<G>(a.out_cx, qid, (G)cx);
            store_as<G>(a.out_cy, qid, (G)cy);
            store_as<G>(a.out_cz, qid, (G)cz);
            if (a.out_tri) a.out_tri[qid] = bt;
        }
        return;
    }
    double nl64;
    const double dot = sign_dot<double>(qx, qy, qz, A, B, C, bt, a.s0, a.s1, a.s2, a.face_n, a.edge_n,
                                        a.vertex_n, nl64);
    store_as<G>(a.out_d, qid, (G)(dot < 0 ? -dist : dist));
    if (a.out_tri) a.out_tri[qid] = bt;
}

// ------------------------------------------------------------------------------------------------
// Per-query search, one wave per queued query.
//
// The queue holds FIX queries whose band has more triangles than the candidate list (word 1: the
// FP32 band) and the queries of packet waves that ran past their leaf budget (word 1: FIX, the
// band of the best found so far; plain, the best so far).  The wave walks the tree deepest level
// first: a step takes up to 64 pending nodes of the deepest non-empty level, one per lane; the
// lane tests both child boxes against the bound, evaluates a passing leaf at once and appends a
// passing internal child to the next level.  Children enter a level only while every deeper level
// is empty, so a level never holds more than 128 entries.  After a step in which some lane
// improved, the bound tightens to the wave's best.  Waves take entries one at a time from a
// counter, since their costs differ by orders of magnitude.
//   FIX (FP64): the search of exact_one: inclusive bound (sqrt(band2) + delta)^2, FP64 distances,
//     tie-break (d2, original index), so the canonical result.
//   plain (FP32): inclusive bound (the triangle that gave it is found again), the packet
//     traversal's FP32 distances, tie-break (d2, original index), the packet's epilogue.  The
//     packet walk keeps the first triangle it meets among exact FP32 ties and neither walk is
//     rounding-safe, so the two may differ on ties and by an ulp, as the plain contract allows.
// ------------------------------------------------------------------------------------------------
constexpr int kCoopBlocks = 1024;  // up to 4 waves each, taking queue entries until none is left


template <typename T>
__device__ __forceinline__ void wave_min3(T &d, int &t, int &k) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const T od = __shfl_xor(d, off);
        const int ot = __shfl_xor(t, off), ok = __shfl_xor(k, off);
        if (od < d || (od == d && ot < t)) d = od, t = ot, k = ok;
    }
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// box(mn, mx, bound): does a child box pass; leaf(k): evaluates leaf position k into the lane's
// best, true if it improved; tighten(bound, wave best).  lev: an LDS stack of coop_stack(nlev)
// ints holding the pending levels one after the other, the deepest on top: only the top level can
// hold 128 entries, every other one keeps at most the 64 its first step left.  False when the walk
// needs a level past nlev - 1 (not a tree of that depth).
__host__ __device__ constexpr int coop_stack(int nlev) { return 64 * (nlev + 1); }

// Waves per block of the per-query kernels: 4 when their stacks fit 64 KiB of LDS, else fewer.
static int coop_waves(int nlev) {
    const int per = (int)sizeof(int) * coop_stack(nlev);
    return std::max(1, std::min(4, 65536 / per));
}

template <typename T, typename Box, typename Leaf, typename Tighten>
__device__ __forceinline__ bool coop_walk(const Node *__restrict__ nodes, const int root, int *__restrict__ lev,
                                          const int nlev, T &bound, T &best, int &bt, int &bk, Box box, Leaf leaf,
                                          Tighten tighten) {
    const int lane = threadIdx.x & 63;
    const uint64_t below = (1ull << lane) - 1ull;
    if (root < 0) {  // one triangle: every lane evaluates it
        leaf(~root);
        return true;
    }
    // pending entries of level L: lane L of cnt_lo (L < 64) or lane L - 64 of cnt_hi
    int cnt_lo = lane == 0 ? 1 : 0, cnt_hi = 0;
    if (lane == 0) lev[0] = root;
    wave_lds_sync();
    int top = 0, sp = 1;  // deepest pending level, stack size (wave-uniform)
    while (top >= 0) {
        const int c = top < 64 ? __builtin_amdgcn_readlane(cnt_lo, top) : __builtin_amdgcn_readlane(cnt_hi, top - 64);
        const int take = min(c, 64);
        if (top < 64)
            cnt_lo = lane == top ? c - take : cnt_lo;
        else
            cnt_hi = lane == top - 64 ? c - take : cnt_hi;
        sp -= take;
        bool p0 = false, p1 = false, imp = false;
        int c0 = 0, c1 = 0;
        if (lane < take) {
            const Node nd = nodes[lev[sp + lane]];
            p0 = box(nd.lmin, nd.lmax, bound);
            p1 = box(nd.rmin, nd.rmax, bound);
            c0 = nd.left, c1 = nd.right;
            // leaf references: a run of leaves, every one evaluated
            if (p0 && c0 < 0) {
                for (int j = 0; j < leaf_count(c0); j++) imp |= leaf(leaf_first(c0) + j);
                p0 = false;
            }
            if (p1 && c1 < 0) {
                for (int j = 0; j < leaf_count(c1); j++) imp |= leaf(leaf_first(c1) + j);
                p1 = false;
            }
        }
        const uint64_t m0 = __ballot(p0), m1 = __ballot(p1);
        const int n = __popcll(m0) + __popcll(m1);
        if (n > 0) {
            if (top + 1 >= nlev || sp + n > coop_stack(nlev)) return false;
            int *dst = lev + sp;
            if (p0) dst[__popcll(m0 & below)] = c0;
            if (p1) dst[__popcll(m0) + __popcll(m1 & below)] = c1;
            sp += n;
            top++;
            if (top < 64)
                cnt_lo = lane == top ? n : cnt_lo;
            else
                cnt_hi = lane == top - 64 ? n : cnt_hi;
        } else {
            const uint64_t h0 = __ballot(cnt_lo > 0), h1 = __ballot(cnt_hi > 0);
            top = h1 ? 127 - __clzll(h1) : (h0 ? 63 - __clzll(h0) : -1);
        }
        if (__ballot(imp) != 0) {
            T wb = best;
            int wt = bt, wk = bk;
            wave_min3(wb, wt, wk);
            tighten(bound, wb);
        }
        wave_lds_sync();
    }
    return true;
}

// FIX entry: the exact FP64 search below band2, result written as exact_one writes it.
template <int QSRC, int OUT>
__device__ bool coop_exact(const QueryArgs &a, const long long qid, const double band2, int *lev, const int nlev) {
    double qx, qy, qz;
    query_coords<float, QSRC>(a, qid, qx, qy, qz);
    double out_init = __builtin_inf();
    if (OUT == OUT_EDF || OUT == OUT_CLOSEST || OUT == OUT_SQEDF) out_init = (double)load_as<float>(a.out_d, qid);
    const double init2 = OUT == OUT_SQEDF ? out_init : out_init * out_init;  // SQEDF seeds are squared
    const double delta = fmax(a.delta64, 1e-13 * fmax(fabs(qx), fmax(fabs(qy), fabs(qz))));
    const double b0 = sqrt(fmin(init2, band2)) + delta;
    double bound = b0 * b0;
    double best = __builtin_inf();
    int bt = INT_MAX, bk = -1;
    const bool ok = coop_walk<double>(
        a.nodes, uniform_i(a.meta->root), lev, nlev, bound, best, bt, bk,
        [&](const float *mn, const float *mx, const double bnd) {
            return pt_box_d2<double>(qx, qy, qz, mn[0], mn[1], mn[2], mx[0], mx[1], mx[2]) <= bnd;
        },
        [&](const int k) {
            const TriF tr = a.tris[k].f;
            double A[3], B[3], C[3];
            tri_verts64<float>(a, tr, A, B, C);
            const double d = pt_tri_d2<double>(qx, qy, qz, A[0], A[1], A[2], B[0], B[1], B[2], C[0], C[1], C[2]);
            if (d < init2 && (d < best || (d == best && tr.orig < bt))) {
                best = d, bt = tr.orig, bk = k;
                return true;
            }
            return false;
        },
        [&](double &bnd, const double wb) {
            if (wb < __builtin_inf()) {
                const double r = sqrt(wb) + delta;
                bnd = fmin(bnd, r * r);
            }
        });
    if (!ok) return false;
    wave_min3(best, bt, bk);
    if ((threadIdx.x & 63) == 0) exact_write<float, OUT>(a, qid, qx, qy, qz, out_init, best, bt, bk);
    return true;
}

// Plain entry: the FP32 minimum at or below bound0 (the FP32 band of the packet wave's best so far),
// written by the packet epilogue.
template <int QSRC, int OUT>
__device__ bool coop_plain(const QueryArgs &a, const long long qid, const float bound0, int *lev, const int nlev) {
    double dx, dy, dz;
    query_coords<float, QSRC>(a, qid, dx, dy, dz);  // FP32 coordinates, widened exactly
    const float qx = (float)dx, qy = (float)dy, qz = (float)dz;
    float out_init = SSDF_INF, init2 = SSDF_INF;
    if (OUT == OUT_EDF || OUT == OUT_CLOSEST) {
        out_init = ((const float *)a.out_d)[qid];
        init2 = out_init * out_init;
    }
    if (OUT == OUT_SQEDF) init2 = ((const float *)a.out_d)[qid];
    const float babs = fmaxf(a.band_abs, 0x1p-18f * fmaxf(fabsf(qx), fmaxf(fabsf(qy), fabsf(qz))));
    float bound = bound0, best = SSDF_INF;
    int bt = INT_MAX, bk = -1;
    const TriP *__restrict__ tris = a.tris;
    const bool ok = coop_walk<float>(
        a.nodes, uniform_i(a.meta->root), lev, nlev, bound, best, bt, bk,
        [&](const float *mn, const float *mx, const float bnd) { return box_d2_fast(qx, qy, qz, mn, mx) <= bnd; },
        [&](const int k) {
            const TriP &tr = tris[k];
            const float d2 = tri_d2_fast(qx, qy, qz, tr);
            const int orig = tr.f.orig;
            if (d2 < init2 && (d2 < best || (d2 == best && orig < bt))) {
                best = d2, bt = orig, bk = k;
                return true;
            }
            return false;
        },
        [&](float &bnd, const float wb) { bnd = fminf(bnd, wb); });
    if (!ok) return false;
    wave_min3(best, bt, bk);
    if ((threadIdx.x & 63) == 0) packet_write<OUT>(a, qid, qx, qy, qz, out_init, init2, babs, bk, best, false, 0.0);
    return true;
}

template <int QSRC, int OUT, bool FIX>
__device__ __forceinline__ void coop_item(const QueryArgs &a, const unsigned *e, int *lev, const int nlev) {
    const long long qid = (long long)(((unsigned long long)e[3] << 32) | e[0]);
    const float b = __uint_as_float(e[1]);
    const bool ok = FIX ? coop_exact<QSRC, OUT>(a, qid, (double)b, lev, nlev)
                        : coop_plain<QSRC, OUT>(a, qid, b, lev, nlev);
    if (!ok && (threadIdx.x & 63) == 0) atomicOr(a.err_flag, 1);
    wave_lds_sync();  // the next item reuses the levels
}

// Next queue entry for the calling wave (the entries differ in cost by orders of magnitude, so
// they are handed out one at a time); n_items or more when the queue is exhausted.  A plain read
// first, so that waves arriving after the end (or at an empty queue: most calls) take no atomic.
__device__ __forceinline__ unsigned coop_next(const QueryArgs &a, const unsigned n_items) {
    unsigned item = 0;
    if ((threadIdx.x & 63) == 0) {
        item = __hip_atomic_load(a.fix_count + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (item < n_items) item = atomicAdd(a.fix_count + 2, 1u);
    }
    return (unsigned)__shfl((int)item, 0);
}

template <int QSRC, int OUT, bool FIX>
__global__ __launch_bounds__(256) void k_coop(QueryArgs a, int nlev) {
    extern __shared__ int lds_levels[];
    int *lev = lds_levels + (threadIdx.x >> 6) * coop_stack(nlev);
    const unsigned n_items = min(*a.fix_count, a.fix_cap);
    for (unsigned item = coop_next(a, n_items); item < n_items; item = coop_next(a, n_items))
        coop_item<QSRC, OUT, FIX>(a, a.fix_list + 4 * (size_t)item, lev, nlev);
}

template <int QSRC, int OUT>
static void launch_coop(const QueryArgs &a, bool fix, int nlev, hipStream_t s) {
    const int w = coop_waves(nlev);
    const size_t lds = sizeof(int) * (size_t)coop_stack(nlev) * w;
    if (fix)
        hipLaunchKernelGGL((k_coop<QSRC, OUT, true>), dim3(kCoopBlocks), dim3(64 * w), lds, s, a, nlev);
    else
        hipLaunchKernelGGL((k_coop<QSRC, OUT, false>), dim3(kCoopBlocks), dim3(64 * w), lds, s, a, nlev);
}

// Every query of the call, one lane each (grid-stride).
template <typename G, int QSRC, int OUT>
__global__ __launch_bounds__(64) void k_query_exact(QueryArgs a) {
    const long long total = QSRC == QSRC_POINTS ? a.npoints : (long long)a.nx * a.ny * a.nz;
    for (long long item = (long long)blockIdx.x * blockDim.x + threadIdx.x; item < total;
         item += (long long)gridDim.x * blockDim.x)
        exact_one<G, QSRC, OUT>(a, item, __builtin_inf());
}

// ------------------------------------------------------------------------------------------------
// Batches
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ void batch_args(const BatchArgs &b, const BatchMesh &bm, QueryArgs &a) {
    a = b.base;
    a.nodes = b.nodes + bm.node_off;
    a.tris = b.tris + bm.tri_off;
    a.meta = b.metas + bm.mesh;
    a.s0 = b.s0 + bm.tri_off, a.s1 = b.s1 + bm.tri_off, a.s2 = b.s2 + bm.tri_off;
    a.sx = b.sx + bm.vert_off, a.sy = b.sy + bm.vert_off, a.sz = b.sz + bm.vert_off;
    a.face_n = b.face_n ? b.face_n + 3 * bm.tri_off : nullptr;
    a.edge_n = b.edge_n ? b.edge_n + 9 * bm.tri_off : nullptr;
    a.vertex_n = b.vertex_n ? b.vertex_n + 3 * bm.vert_off : nullptr;
    a.qx = b.qx ? b.qx + bm.out_off : nullptr;
    a.qy = b.qy ? b.qy + bm.out_off : nullptr;
    a.qz = b.qz ? b.qz + bm.out_off : nullptr;
    a.npoints = bm.npoints;
    a.nx = bm.nx, a.ny = bm.ny, a.nz = bm.nz, a.nbx = bm.nbx, a.nby = bm.nby, a.nbz = bm.nbz;
    a.brick = bm.brick;
    if (bm.heavy >= 0) a.heavy_iters = bm.heavy;
#pragma unroll
    for (int d = 0; d < 3; d++) a.gmin[d] = bm.gmin[d], a.gh[d] = bm.gh[d];
    a.out_d = b.out_d + bm.out_off;
    a.out_tri = b.out_tri ? b.out_tri + bm.out_off : nullptr;
    a.out_cx = b.out_cx ? b.out_cx + bm.out_off : nullptr;
    a.out_cy = b.out_cy ? b.out_cy + bm.out_off : nullptr;
    a.out_cz = b.out_cz ? b.out_cz + bm.out_off : nullptr;
    a.band_abs = bm.band_abs, a.sign_eps = bm.sign_eps, a.delta64 = bm.delta64;
    a.fix_tag = (unsigned)bm.mesh;
}

// Persistent: every wave takes `chunk` units at a time from a work queue until all are done.  The
// units are split into 8 contiguous ranges with a counter each; block b starts on range b % 8 (the
// XCD the hardware dispatches it to, so each XCD's L2 holds mostly its range's meshes) and moves
// on to the next ranges when its own is exhausted.  With an order, position p of a range is work
// unit order[p] (a permutation within the range: the batch puts the units its previous query
// measured costliest first), and each wave records its unit's cycles in unit_cost.
template <int STACK, int QSRC, int OUT, bool FIX>
// Occupancy, measured on the 64-mesh batch (FP32 + fixup, resident query): 8 waves per SIMD (64
// VGPRs, the rest spilled outside the traversal loop) 76.5 ms, 7 waves 79.1 ms, 6 waves 83.2 ms,
// unconstrained (92 VGPRs, 5 waves) 89.5 ms.  Plain runs best at 8 as well.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_batch_packet(const Node *__restrict__ nodes, const TriP *__restrict__ tris,
                                                      const BatchMesh *__restrict__ meshes, BatchArgs b) {
    __shared__ int cand[FIX ? kCand : 1][256];
    const int lane = threadIdx.x & 63;
    const unsigned U = b.total_units;
    const unsigned x0 = blockIdx.x & 7u;
    for (unsigned k = 0; k < 8; k++) {
        const unsigned x = (x0 + k) & 7u;
        const unsigned lo = (unsigned)(((unsigned long long)U * x) >> 3);
        const unsigned hi = (unsigned)(((unsigned long long)U * (x + 1)) >> 3);
        for (;;) {
            unsigned grab = 0;
            if (lane == 0) grab = atomicAdd(&b.queue[64 * x], b.chunk);
            grab = (unsigned)uniform_i((int)__shfl((int)grab, 0));
            const unsigned u0 = lo + grab;
            if (grab >= hi - lo) break;
            const unsigned u1 = min(u0 + b.chunk, hi);
            for (unsigned pos = u0; pos < u1; pos++) {
                const unsigned u = b.order ? b.order[pos] : pos;
                int m0 = 0, m1 = b.nmeshes;
                while (m1 - m0 > 1) {
                    const int mid = (m0 + m1) >> 1;
                    if (meshes[mid].unit_start <= u)
                        m0 = mid;
                    else
                        m1 = mid;
                }
                const BatchMesh &bm = meshes[m0];
                QueryArgs a;
                batch_args(b, bm, a);
                const unsigned long long t0 = (b.cost || b.unit_cost) ? clock64() : 0ull;
                packet_wave<STACK, QSRC, OUT, FIX>(nodes + bm.node_off, tris + bm.tri_off, a, u - bm.unit_start,
                                                   cand);
                if (lane == 0 && (b.cost || b.unit_cost)) {
                    const unsigned long long dt = clock64() - t0;
                    if (b.cost) atomicAdd(b.cost + bm.mesh, dt);
                    if (b.unit_cost) b.unit_cost[u] = (unsigned)min(dt, 0xffffffffull);
                }
            }
        }
    }
}

template <int QSRC, int OUT, bool FIX>
__global__ __launch_bounds__(256) void k_batch_coop(const BatchMesh *__restrict__ meshes, BatchArgs b, int nlev) {
    extern __shared__ int lds_levels[];
    int *lev = lds_levels + (threadIdx.x >> 6) * coop_stack(nlev);
    const unsigned n_items = min(*b.base.fix_count, b.base.fix_cap);
    for (unsigned item = coop_next(b.base, n_items); item < n_items; item = coop_next(b.base, n_items)) {
        const unsigned *e = b.base.fix_list + 4 * (size_t)item;
        const int mi = (int)e[2];
        // descriptors are ordered by unit, not by mesh: find the mesh's entry
        int j = 0;
        while (j < b.nmeshes - 1 && meshes[j].mesh != mi) j++;
        QueryArgs a;
        batch_args(b, meshes[j], a);
        const unsigned long long t0 = b.cost ? clock64() : 0ull;
        coop_item<QSRC, OUT, FIX>(a, e, lev, nlev);
        if (b.cost && (threadIdx.x & 63) == 0) atomicAdd(b.cost + mi, (unsigned long long)(clock64() - t0));
    }
}

__global__ void k_batch_reset(unsigned *queue, unsigned *status) {
    if (threadIdx.x < 8) queue[64 * threadIdx.x] = 0u;
    if (threadIdx.x < 4) status[threadIdx.x] = 0u;
}

template <int STACK, int QSRC, int OUT>
static void launch_batch_packet(const BatchArgs &b, const BatchMesh *meshes, bool fix, unsigned nblocks,
                                hipStream_t s) {
    if (fix)
        hipLaunchKernelGGL((k_batch_packet<STACK, QSRC, OUT, true>), dim3(nblocks), dim3(256), 0, s, b.nodes, b.tris,
                           meshes, b);
    else
        hipLaunchKernelGGL((k_batch_packet<STACK, QSRC, OUT, false>), dim3(nblocks), dim3(256), 0, s, b.nodes,
                           b.tris, meshes, b);
    const int w = coop_waves(b.nlev);
    const size_t lds = sizeof(int) * (size_t)coop_stack(b.nlev) * w;
    if (fix)
        hipLaunchKernelGGL((k_batch_coop<QSRC, OUT, true>), dim3(kCoopBlocks), dim3(64 * w), lds, s, meshes, b,
                           b.nlev);
    else
        hipLaunchKernelGGL((k_batch_coop<QSRC, OUT, false>), dim3(kCoopBlocks), dim3(64 * w), lds, s, meshes, b,
                           b.nlev);
}

template <int STACK, int QSRC>
static int launch_batch_out(const BatchArgs &b, const BatchMesh *meshes, int out, bool fix, unsigned nblocks,
                            hipStream_t s) {
    if (out == OUT_EDF)
        launch_batch_packet<STACK, QSRC, OUT_EDF>(b, meshes, fix, nblocks, s);
    else if (out == OUT_SDF)
        launch_batch_packet<STACK, QSRC, OUT_SDF>(b, meshes, fix, nblocks, s);
    else if (out == OUT_SQEDF)
        launch_batch_packet<STACK, QSRC, OUT_SQEDF>(b, meshes, fix, nblocks, s);
    else if (QSRC == QSRC_POINTS)
        launch_batch_packet<STACK, QSRC_POINTS, OUT_CLOSEST>(b, meshes, fix, nblocks, s);
    else
        return SSDF_ERR_UNSUPPORTED;
    return SSDF_OK;
}

int launch_batch(const BatchArgs &b, const BatchMesh *meshes, int qsrc, int out, bool fix, int stack,
                 unsigned nblocks, hipStream_t stream) {
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, stream, b.queue, b.base.fix_count);
    int rc;
    if (stack == 64)
        rc = qsrc == QSRC_POINTS ? launch_batch_out<64, QSRC_POINTS>(b, meshes, out, fix, nblocks, stream)
                                 : launch_batch_out<64, QSRC_GRID>(b, meshes, out, fix, nblocks, stream);
    else
        rc = qsrc == QSRC_POINTS ? launch_batch_out<128, QSRC_POINTS>(b, meshes, out, fix, nblocks, stream)
                                 : launch_batch_out<128, QSRC_GRID>(b, meshes, out, fix, nblocks, stream);
    if (rc) {
        set_error("batch: closest-point output needs point queries");
        return rc;
    }
    SSDF_HIP_TRY(hipGetLastError());
    return SSDF_OK;
}

// ------------------------------------------------------------------------------------------------
// Morton order of scattered query points
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ __launch_bounds__(256) void k_point_keys(long long n, const float *__restrict__ x,
                                                    const float *__restrict__ y, const float *__restrict__ z,
                                                    const MeshMeta *__restrict__ meta, uint32_t *__restrict__ keys,
                                                    uint32_t *__restrict__ vals) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float p[3] = {x[i], y[i], z[i]};
    uint32_t q[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const float lo = meta->root_box[d], hi = meta->root_box[3 + d];
        const float ext = hi - lo;
        float u = ext > 0.0f ? (p[d] - lo) / ext : 0.0f;
        u = fminf(fmaxf(u, 0.0f), 1.0f);  // outside the box (or NaN): clamped
        q[d] = min((uint32_t)(u * 1024.0f), 1023u);
    }
    keys[i] = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
    vals[i] = (uint32_t)i;
}

int sort_points(const MeshMeta *meta, long long n, const float *x, const float *y, const float *z, void *ws,
                const unsigned **perm, hipStream_t stream) {
    const size_t nn = (size_t)n;
    uint32_t *keys = (uint32_t *)ws, *keys_alt = keys + nn, *vals = keys_alt + nn, *vals_alt = vals + nn;
    uint32_t *scratch = vals_alt + nn;
    hipLaunchKernelGGL(k_point_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, n, x, y, z, meta, keys,
                       vals);
    uint32_t *sk = nullptr, *sv = nullptr;
    SSDF_HIP_TRY(radix_sort_pairs<uint32_t>(keys, keys_alt, vals, vals_alt, nn, 30, scratch, stream, &sk, &sv));
    *perm = sv;
    return SSDF_OK;
}

// ------------------------------------------------------------------------------------------------
// Host dispatch
// ------------------------------------------------------------------------------------------------
// Compute units of the current device (cached per device).
unsigned device_cus() {
    static std::atomic<unsigned> cus[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256u;
    unsigned c = cus[dev].load(std::memory_order_relaxed);
    if (!c) {
        int n = 0;
        c = (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0) ? (unsigned)n
                                                                                                          : 256u;
        cus[dev].store(c, std::memory_order_relaxed);
    }
    return c;
}
// Resident blocks per compute unit of k_query_packet: 8 waves per SIMD, 4 waves a block.
constexpr unsigned kBlocksPerCu = 8;

template <int STACK, int QSRC, int OUT, bool FIX>
static void launch_packet(QueryArgs &a, unsigned nblocks, hipStream_t s) {
    a.nblocks = nblocks;
    hipLaunchKernelGGL((k_query_packet<STACK, QSRC, OUT, FIX>), dim3(nblocks), dim3(256), 0, s, a.nodes, a.tris, a);
}

template <int STACK, int QSRC>
static int dispatch_packet_out(QueryArgs &a, int out, bool fix, unsigned nblocks, hipStream_t s) {
    if (out == OUT_EDF) {
        fix ? launch_packet<STACK, QSRC, OUT_EDF, true>(a, nblocks, s)
            : launch_packet<STACK, QSRC, OUT_EDF, false>(a, nblocks, s);
    } else if (out == OUT_SDF) {
        fix ? launch_packet<STACK, QSRC, OUT_SDF, true>(a, nblocks, s)
            : launch_packet<STACK, QSRC, OUT_SDF, false>(a, nblocks, s);
    } else if (out == OUT_SQEDF) {
        fix ? launch_packet<STACK, QSRC, OUT_SQEDF, true>(a, nblocks, s)
            : launch_packet<STACK, QSRC, OUT_SQEDF, false>(a, nblocks, s);
    } else {
        if (QSRC != QSRC_POINTS) return SSDF_ERR_UNSUPPORTED;
        fix ? launch_packet<STACK, QSRC_POINTS, OUT_CLOSEST, true>(a, nblocks, s)
            : launch_packet<STACK, QSRC_POINTS, OUT_CLOSEST, false>(a, nblocks, s);
    }
    return SSDF_OK;
}

template <typename G, int QSRC>
static int dispatch_exact_out(QueryArgs &a, int out, unsigned nblocks, hipStream_t s) {
    if (out == OUT_EDF)
        hipLaunchKernelGGL((k_query_exact<G, QSRC, OUT_EDF>), dim3(nblocks), dim3(64), 0, s, a);
    else if (out == OUT_SDF)
        hipLaunchKernelGGL((k_query_exact<G, QSRC, OUT_SDF>), dim3(nblocks), dim3(64), 0, s, a);
    else if (out == OUT_SQEDF)
        hipLaunchKernelGGL((k_query_exact<G, QSRC, OUT_SQEDF>), dim3(nblocks), dim3(64), 0, s, a);
    else
        hipLaunchKernelGGL((k_query_exact<G, QSRC, OUT_CLOSEST>), dim3(nblocks), dim3(64), 0, s, a);
    return SSDF_OK;
}

int launch_query(const DevMesh &m, QueryArgs &a, int qsrc, int out, unsigned flags, int out_is_double,
                 int q_is_double, hipStream_t stream) {
    if (m.nsel <= 0) return SSDF_OK;
    if (out == OUT_SDF && !m.has_normals) {
        set_error("signed queries need the mesh's pseudonormals");
        return SSDF_ERR_INVALID;
    }
    a.nodes = m.nodes;
    a.tris = m.tris;
    a.meta = m.meta;
    a.s0 = m.s0, a.s1 = m.s1, a.s2 = m.s2;
    a.sx = m.sx, a.sy = m.sy, a.sz = m.sz;
    a.face_n = m.face_n, a.edge_n = m.edge_n, a.vertex_n = m.vertex_n;
    a.max_iters = 4 * (long long)m.nsel + 64;
    const long long nq = qsrc == QSRC_POINTS ? a.npoints : (long long)a.nx * a.ny * a.nz;
    if (nq <= 0) return SSDF_OK;
    if (m.dtype == 1) {
        // FP64 mesh: FP64 queries and outputs, exact traversal
        if (!(out_is_double && q_is_double)) {
            set_error("FP64 meshes take FP64 queries and FP64 outputs");
            return SSDF_ERR_UNSUPPORTED;
        }
        const unsigned nb = (unsigned)std::min<long long>((nq + 63) / 64, 1 << 20);
        const int rc = qsrc == QSRC_POINTS ? dispatch_exact_out<double, QSRC_POINTS>(a, out, nb, stream)
                                           : dispatch_exact_out<double, QSRC_GRID>(a, out, nb, stream);
        SSDF_HIP_TRY(hipGetLastError());
        return rc;
    }
    if (out_is_double || q_is_double) {
        set_error("FP32 meshes take FP32 queries and FP32 outputs");
        return SSDF_ERR_UNSUPPORTED;
    }
    const bool fix = (flags & QF_FIXUP) != 0;
    a.flags = flags;
    const int depth = m.hmeta.maxdepth;
    unsigned nwaves;
    if (qsrc == QSRC_POINTS) {
        nwaves = (unsigned)((nq + 63) / 64);
    } else {
        unsigned long long w = 0;
        grid_bricks(a.brick, a.nx, a.ny, a.nz, a.nbx, a.nby, a.nbz, w);
        if (w > 0xffffffffull - kWaves) {
            set_error("grid query: too many nodes for one launch");
            return SSDF_ERR_INVALID;
        }
        nwaves = (unsigned)w;
    }
    unsigned nblocks = (nwaves + kWaves - 1) / kWaves;
    // More waves than the GPU holds at once (8 per SIMD): a persistent launch of one resident set,
    // the waves handed out per XCD with stealing, instead of one block per 4 waves dispatched in
    // order (which leaves XCDs whose share of the grid is cheaper idle).
    const unsigned resident = device_cus() * kBlocksPerCu;
    a.total_waves = nwaves;
    a.persistent = nblocks > resident && a.wqueue ? 1u : 0u;
    if (a.persistent) nblocks = resident;
    int rc;
    if (depth <= 64 && m.nsel - 1 < kNarrowNodes) {
        rc = qsrc == QSRC_POINTS ? dispatch_packet_out<64, QSRC_POINTS>(a, out, fix, nblocks, stream)
                                 : dispatch_packet_out<64, QSRC_GRID>(a, out, fix, nblocks, stream);
    } else if (depth <= 128) {
        rc = qsrc == QSRC_POINTS ? dispatch_packet_out<128, QSRC_POINTS>(a, out, fix, nblocks, stream)
                                 : dispatch_packet_out<128, QSRC_GRID>(a, out, fix, nblocks, stream);
    } else {
        set_error("LBVH deeper than 128 levels");
        return SSDF_ERR_DEPTH;
    }
    if (rc) return rc;
    if (fix || a.heavy_iters > 0) {
        // queued queries: one wave each, grid-stride over the device-side count
        const int nlev = std::max(m.hmeta.maxdepth, 1);
        if (qsrc == QSRC_POINTS) {
            if (out == OUT_EDF)
                launch_coop<QSRC_POINTS, OUT_EDF>(a, fix, nlev, stream);
            else if (out == OUT_SDF)
                launch_coop<QSRC_POINTS, OUT_SDF>(a, fix, nlev, stream);
            else if (out == OUT_SQEDF)
                launch_coop<QSRC_POINTS, OUT_SQEDF>(a, fix, nlev, stream);
            else
                launch_coop<QSRC_POINTS, OUT_CLOSEST>(a, fix, nlev, stream);
        } else {
            if (out == OUT_EDF)
                launch_coop<QSRC_GRID, OUT_EDF>(a, fix, nlev, stream);
            else if (out == OUT_SQEDF)
                launch_coop<QSRC_GRID, OUT_SQEDF>(a, fix, nlev, stream);
            else
                launch_coop<QSRC_GRID, OUT_SDF>(a, fix, nlev, stream);
        }
    }
    SSDF_HIP_TRY(hipGetLastError());
    return rc;
}

}  // namespace ssdf_amd
