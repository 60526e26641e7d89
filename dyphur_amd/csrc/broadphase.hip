// broadphase.hip — LBVH broadphase for gfx950 (core/src/broadphase.cpp:101-365).
//
// Pipeline (all on one stream, no host synchronisation):
//   1. morton_sort     one workgroup: 30-bit Morton codes of body centres, then a
//                      stable 4-pass LSD radix sort of (code, body) — the same
//                      order as the reference's sort of (code << 32 | body).
//   2. karras_build    one thread per internal node: children, parents, leaf
//                      ranges and the depth-first "escape" links that make the
//                      traversal stackless.
//   3. leaf_refit      one workgroup per 1024 leaves: fp16 leaf boxes, then a
//                      counter-gated bottom-up merge inside the chunk (LDS
//                      counters and the chunk's topology in LDS); refit_top,
//                      one workgroup, then merges the few nodes whose ranges
//                      cross chunks (the launch boundary publishes the
//                      chunks' nodes to it).
//   4. traverse_team   a team of 16 lanes per leaf (Morton order), 256-thread
//                      workgroups with contiguous leaf runs per XCD: the team
//                      expands up to 16 nodes of its LDS stack per step,
//                      collects the partners b > a in LDS and ranks them into
//                      the row (16 smallest) and overflow buffer, ascending.
//                      (traverse_count: one lane per leaf, kept as a variant.)
//   5. scan of counts  row offsets per body a; the total is the true pair count.
//   6. emit_coop       one wavefront per 64 bodies writes their contiguous
//                      output range with one pair per lane per store; rows
//                      longer than 16 + 128 re-walk the tree for the rest.
// The output is the reference's pair set in ascending (a, b) order, so
// sort_pairs has nothing left to do.
#include <cstdlib>

#include "dy_internal.hpp"

namespace dy {

constexpr uint32_t kRow = 16;            // partners kept in registers per walk
constexpr uint32_t kOvf = 128;           // further partners kept per body in an overflow buffer
constexpr uint32_t kChunk = 1024;        // leaves per refit workgroup
constexpr uint32_t kMortonMax = 1u << 30;
constexpr uint32_t kXcds = 8;           // XCDs of an MI355X; workgroups are dealt to them round-robin
// A Karras tree over distinct 64-bit keys is at most 65 levels deep; walks
// that exceed this bound flag an internal error instead of looping.
constexpr int kMaxHeight = 96;

// Traversal record, 32 B: fp16 box, subtree max body index, first child
// (internal) or body (leaf), escape link.
struct alignas(16) TNode {
    uint32_t b01, b23, b45;  // (mn_x | mn_y << 16), (mn_z | mx_x << 16), (mx_y | mx_z << 16)
    uint32_t maxbody;
    int32_t child;
    int32_t escape;
    uint32_t pad0, pad1;
};

// Walk record of internal node p, 48 B: both children's fp16 boxes, subtree
// max body indices and node numbers, so a walk tests the children of the
// node it loads and never loads a leaf.
struct alignas(16) TPair {
    uint32_t lb01, lb23, lb45, lmax;
    uint32_t rb01, rb23, rb45, rmax;
    int32_t left, right;
    uint32_t pad0, pad1;
};

struct Box6 { float mn[3], mx[3]; };

__device__ __forceinline__ Box6 unpack_box(uint32_t b01, uint32_t b23, uint32_t b45) {
    Box6 b;
    b.mn[0] = f16_to_f32(static_cast<uint16_t>(b01 & 0xffffu));
    b.mn[1] = f16_to_f32(static_cast<uint16_t>(b01 >> 16));
    b.mn[2] = f16_to_f32(static_cast<uint16_t>(b23 & 0xffffu));
    b.mx[0] = f16_to_f32(static_cast<uint16_t>(b23 >> 16));
    b.mx[1] = f16_to_f32(static_cast<uint16_t>(b45 & 0xffffu));
    b.mx[2] = f16_to_f32(static_cast<uint16_t>(b45 >> 16));
    return b;
}

// Reference overlap test (broadphase.cpp:322-325): touching counts.
__device__ __forceinline__ bool overlaps(const Box6 &n, const Box6 &q) {
    return !(n.mn[0] > q.mx[0] || n.mx[0] < q.mn[0] || n.mn[1] > q.mx[1] || n.mx[1] < q.mn[1] ||
             n.mn[2] > q.mx[2] || n.mx[2] < q.mn[2]);
}

// Union of two packed fp16 boxes (exact: fmin/fmax of representable values).
__device__ __forceinline__ void merge_packed(const uint32_t a[3], const uint32_t b[3], uint32_t o[3]) {
    uint16_t ha[6] = {static_cast<uint16_t>(a[0] & 0xffffu), static_cast<uint16_t>(a[0] >> 16),
                      static_cast<uint16_t>(a[1] & 0xffffu), static_cast<uint16_t>(a[1] >> 16),
                      static_cast<uint16_t>(a[2] & 0xffffu), static_cast<uint16_t>(a[2] >> 16)};
    uint16_t hb[6] = {static_cast<uint16_t>(b[0] & 0xffffu), static_cast<uint16_t>(b[0] >> 16),
                      static_cast<uint16_t>(b[1] & 0xffffu), static_cast<uint16_t>(b[1] >> 16),
                      static_cast<uint16_t>(b[2] & 0xffffu), static_cast<uint16_t>(b[2] >> 16)};
    uint16_t r[6];
#pragma unroll
    for (int c = 0; c < 3; ++c) r[c] = f32_to_f16(fminf(f16_to_f32(ha[c]), f16_to_f32(hb[c])));
#pragma unroll
    for (int c = 3; c < 6; ++c) r[c] = f32_to_f16(fmaxf(f16_to_f32(ha[c]), f16_to_f32(hb[c])));
    o[0] = r[0] | (static_cast<uint32_t>(r[1]) << 16);
    o[1] = r[2] | (static_cast<uint32_t>(r[3]) << 16);
    o[2] = r[4] | (static_cast<uint32_t>(r[5]) << 16);
}

__device__ __forceinline__ void store_pair(TPair *tp, int par, const uint32_t lb[3], uint32_t lmax,
                                           const uint32_t rb[3], uint32_t rmax, int l, int r) {
    uint4 *w = reinterpret_cast<uint4 *>(&tp[par]);
    w[0] = make_uint4(lb[0], lb[1], lb[2], lmax);
    w[1] = make_uint4(rb[0], rb[1], rb[2], rmax);
    w[2] = make_uint4(static_cast<uint32_t>(l), static_cast<uint32_t>(r), 0u, 0u);
}

__device__ __forceinline__ void store_node(dy_node_box *nb, TNode *tn, int node, const uint32_t p[3],
                                           uint32_t maxbody) {
    // Reference NodeBox layout (pads are zero) and the traversal record.
    uint4 v = make_uint4(p[0], p[1], p[2], 0u);
    *reinterpret_cast<uint4 *>(&nb[node]) = v;
    tn[node].b01 = p[0];
    tn[node].b23 = p[1];
    tn[node].b45 = p[2];
    tn[node].maxbody = maxbody;
}

// ---------------------------------------------------------------------------
// 1. Morton codes + stable radix sort, one workgroup
// ---------------------------------------------------------------------------
constexpr uint32_t kMsThreads = 1024;
constexpr uint32_t kMsWaves = kMsThreads / 64;
constexpr uint32_t kMsItems = 4;
constexpr uint32_t kMsTile = kMsThreads * kMsItems;

__global__ __launch_bounds__(kMsThreads) void morton_sort(dy_body_view bv, dy_aabb bd, uint32_t n,
                                                          uint32_t *code0, uint32_t *idx0, uint32_t *code1,
                                                          uint32_t *idx1, uint64_t *morton64) {
    __shared__ uint32_t hist[256];
    __shared__ uint32_t base[256];
    __shared__ uint32_t wcnt[kMsWaves][256];
    const uint32_t t = threadIdx.x, lane = lane_id(), wave = t >> 6;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (uint32_t i = t; i < n; i += kMsThreads) {
        const uint32_t qx = quantize10(bv.pos_x[i], bd.min_x, bd.max_x);
        const uint32_t qy = quantize10(bv.pos_y[i], bd.min_y, bd.max_y);
        const uint32_t qz = quantize10(bv.pos_z[i], bd.min_z, bd.max_z);
        code0[i] = morton3(qx, qy, qz);
        idx0[i] = i;
    }
    __syncthreads();
    uint32_t *sk = code0, *sv = idx0, *dk = code1, *dv = idx1;
    for (int bit = 0; bit < 30; bit += 8) {
        const uint32_t mask = (bit == 24) ? 0x3fu : 0xffu;
        if (t < 256) hist[t] = 0;
        __syncthreads();
        for (uint32_t i = t; i < n; i += kMsThreads) atomicAdd(&hist[(sk[i] >> bit) & mask], 1u);
        __syncthreads();
        if (wave == 0) {
            uint32_t h[4], s = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) { h[j] = hist[lane * 4 + j]; s += h[j]; }
            uint32_t run = wave_inclusive_scan(s) - s;
#pragma unroll
            for (int j = 0; j < 4; ++j) { base[lane * 4 + j] = run; run += h[j]; }
        }
        __syncthreads();
        for (uint32_t tile = 0; tile < n; tile += kMsTile) {
            for (uint32_t i = t; i < kMsWaves * 256; i += kMsThreads) (&wcnt[0][0])[i] = 0;
            __syncthreads();
            uint32_t kk[kMsItems], vv[kMsItems], dd[kMsItems], rr[kMsItems];
#pragma unroll
            for (uint32_t j = 0; j < kMsItems; ++j) {
                const uint32_t idx = tile + wave * (64 * kMsItems) + j * 64 + lane;
                const bool valid = idx < n;
                kk[j] = valid ? sk[idx] : 0u;
                vv[j] = valid ? sv[idx] : 0u;
                const uint32_t d = valid ? ((kk[j] >> bit) & mask) : 256u;
                uint64_t peers = ~0ull;
#pragma unroll
                for (int b = 0; b < 9; ++b) {
                    const bool set = (d >> b) & 1u;
                    const uint64_t bal = __ballot(set);
                    peers &= set ? bal : ~bal;
                }
                const uint32_t pre = static_cast<uint32_t>(__popcll(peers & lt_mask));
                const uint32_t c = valid ? wcnt[wave][valid ? d : 0u] : 0u;
                rr[j] = c + pre;
                dd[j] = d;
                if (valid && pre == 0) wcnt[wave][d] = c + static_cast<uint32_t>(__popcll(peers));
                __builtin_amdgcn_wave_barrier();
            }
            __syncthreads();
            if (t < 256) {
                uint32_t run = base[t];
                for (uint32_t w = 0; w < kMsWaves; ++w) {
                    const uint32_t c = wcnt[w][t];
                    wcnt[w][t] = run;
                    run += c;
                }
                base[t] = run;
            }
            __syncthreads();
#pragma unroll
            for (uint32_t j = 0; j < kMsItems; ++j) {
                if (dd[j] < 256u) {
                    const uint32_t pos = wcnt[wave][dd[j]] + rr[j];
                    dk[pos] = kk[j];
                    dv[pos] = vv[j];
                }
            }
            __syncthreads();
        }
        uint32_t *tk = sk; sk = dk; dk = tk;
        uint32_t *tv = sv; sv = dv; dv = tv;
        __syncthreads();
    }
    // Four passes: the sorted data is back in (code0, idx0).
    for (uint32_t i = t; i < n; i += kMsThreads)
        morton64[i] = (static_cast<uint64_t>(code0[i]) << 32) | idx0[i];
}

// Multi-workgroup form of the same sort for larger n: one cooperative launch
// whose workgroups on one XCD do the work (the others check in and leave),
// four 8-bit LSD passes over 1024-key tiles.  Workgroup b owns a contiguous
// run of tiles.  Pass p:
//   A. each workgroup counts the digits of its tiles' keys (pass p's input
//      order) into hist[tile][digit] (pass 0: while computing the keys);
//   B. after a barrier, every workgroup reads the counts of all tiles (digit
//      d's keys in the tiles before its own and in total; a scan of the totals
//      gives the digit bases), ranks its keys within each tile (wave ballots,
//      waves in index order: stable) and scatters them;
// and a barrier between the passes: seven barriers in all.  The participants
// share the XCD's L2, so the barriers need no cache maintenance: whatever the
// kernel itself wrote (keys, counts) is read with agent-scope loads, which
// skip the vector L1 and are served by that L2.  The last pass also writes
// the 64-bit keys (code << 32 | body).
constexpr uint32_t kSgThreads = 256;
constexpr uint32_t kSortXcc = 0;
constexpr uint32_t kSortMaxBlocks = 64;  // default cap on participants (more pay more per barrier)
constexpr uint32_t kSgWaves = kSgThreads / 64;
constexpr uint32_t kSgItems = 4;
constexpr uint32_t kSgTile = kSgThreads * kSgItems;
constexpr uint32_t kErrSortBarrier = 8u;

__device__ __forceinline__ uint32_t sort_digit(uint32_t key, uint32_t pass) {
    return (key >> (8u * pass)) & (pass == 3u ? 0x3fu : 0xffu);
}

__global__ __launch_bounds__(kSgThreads) void morton_sort_grid(dy_body_view bv, dy_aabb bd, uint32_t n,
                                                               uint32_t n_tiles, uint32_t *code0, uint32_t *idx0,
                                                               uint32_t *code1, uint32_t *idx1, uint32_t *hist,
                                                               uint64_t *morton64, uint32_t *bar, uint32_t *reg,
                                                               uint32_t *err) {
    __shared__ uint32_t lh[256];
    __shared__ uint32_t dbase[256];
    __shared__ uint32_t wcnt[kSgWaves][256];
    __shared__ uint32_t wsum[17];
    uint32_t gen = 0, bidx = 0, nblk = 0;
    if (!xcd_checkin(kSortXcc, reg, err, kErrSortBarrier, bidx, nblk)) return;
    const uint32_t t = threadIdx.x, lane = lane_id(), wave = t >> 6;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    // balanced contiguous runs: tile counts differ by at most one
    const uint32_t t0 = static_cast<uint32_t>(static_cast<uint64_t>(bidx) * n_tiles / nblk);
    const uint32_t t1 = static_cast<uint32_t>(static_cast<uint64_t>(bidx + 1u) * n_tiles / nblk);
    for (uint32_t tile = t0; tile < t1; ++tile) {
        lh[t] = 0;
        __syncthreads();
#pragma unroll
        for (uint32_t j = 0; j < kSgItems; ++j) {
            const uint32_t i = tile * kSgTile + j * kSgThreads + t;
            if (i < n) {
                const uint32_t qx = quantize10(bv.pos_x[i], bd.min_x, bd.max_x);
                const uint32_t qy = quantize10(bv.pos_y[i], bd.min_y, bd.max_y);
                const uint32_t qz = quantize10(bv.pos_z[i], bd.min_z, bd.max_z);
                const uint32_t code = morton3(qx, qy, qz);
                code0[i] = code;
                idx0[i] = i;
                atomicAdd(&lh[sort_digit(code, 0u)], 1u);
            }
        }
        __syncthreads();
        hist[static_cast<size_t>(tile) * 256u + t] = lh[t];
        __syncthreads();
    }
    if (!xcd_barrier(bar, gen, nblk, err, kErrSortBarrier)) return;
    uint32_t *sk = code0, *sv = idx0, *dk = code1, *dv = idx1;
    for (uint32_t pass = 0; pass < 4; ++pass) {
        uint32_t *h = hist;  // read in B before the barrier that ends the pass, rewritten in A after it
        if (pass > 0) {  // the own tiles' counts
            for (uint32_t tile = t0; tile < t1; ++tile) {
                lh[t] = 0;
                __syncthreads();
#pragma unroll
                for (uint32_t j = 0; j < kSgItems; ++j) {
                    const uint32_t idx = tile * kSgTile + j * kSgThreads + t;
                    if (idx < n) atomicAdd(&lh[sort_digit(ld_agent(&sk[idx]), pass)], 1u);
                }
                __syncthreads();
                h[static_cast<size_t>(tile) * 256u + t] = lh[t];
                __syncthreads();
            }
            if (!xcd_barrier(bar, gen, nblk, err, kErrSortBarrier)) return;
        }
        // digit t: keys in the tiles before t0 and in all tiles
        uint32_t before = 0, total = 0;
        for (uint32_t tt = 0; tt < n_tiles; tt += 8u) {
            uint32_t v[8];
#pragma unroll
            for (uint32_t k = 0; k < 8u; ++k) v[k] = tt + k < n_tiles ? ld_agent(&h[(tt + k) * 256u + t]) : 0u;
#pragma unroll
            for (uint32_t k = 0; k < 8u; ++k) {
                total += v[k];
                before += tt + k < t0 ? v[k] : 0u;
            }
        }
        {
            uint32_t all = 0;
            dbase[t] = block_exclusive_scan(total, wsum, all) + before;
        }
        for (uint32_t tile = t0; tile < t1; ++tile) {
            for (uint32_t i = t; i < kSgWaves * 256u; i += kSgThreads) (&wcnt[0][0])[i] = 0;
            __syncthreads();
            uint32_t kk[kSgItems], vv[kSgItems], dd[kSgItems], rr[kSgItems];
#pragma unroll
            for (uint32_t j = 0; j < kSgItems; ++j) {
                const uint32_t idx = tile * kSgTile + wave * (64 * kSgItems) + j * 64 + lane;
                const bool valid = idx < n;
                kk[j] = valid ? ld_agent(&sk[idx]) : 0u;
                vv[j] = valid ? ld_agent(&sv[idx]) : 0u;
                const uint32_t d = valid ? sort_digit(kk[j], pass) : 256u;
                uint64_t peers = ~0ull;
#pragma unroll
                for (int b = 0; b < 9; ++b) {
                    const bool set = (d >> b) & 1u;
                    const uint64_t bal = __ballot(set);
                    peers &= set ? bal : ~bal;
                }
                const uint32_t pre = static_cast<uint32_t>(__popcll(peers & lt_mask));
                const uint32_t c = valid ? wcnt[wave][valid ? d : 0u] : 0u;
                rr[j] = c + pre;
                dd[j] = d;
                if (valid && pre == 0) wcnt[wave][d] = c + static_cast<uint32_t>(__popcll(peers));
                __builtin_amdgcn_wave_barrier();
            }
            __syncthreads();
            uint32_t in_tile = 0;  // this tile's keys with digit t
            for (uint32_t w = 0; w < kSgWaves; ++w) {
                const uint32_t c = wcnt[w][t];
                wcnt[w][t] = in_tile;
                in_tile += c;
            }
            __syncthreads();
#pragma unroll
            for (uint32_t j = 0; j < kSgItems; ++j) {
                if (dd[j] < 256u) {
                    const uint32_t pos = dbase[dd[j]] + wcnt[wave][dd[j]] + rr[j];
                    if (pos >= n) {  // inconsistent counts: flag instead of writing out of range
                        atomicOr(err, kErrSortBarrier);
                        continue;
                    }
                    dk[pos] = kk[j];
                    dv[pos] = vv[j];
                    if (pass == 3u) morton64[pos] = (static_cast<uint64_t>(kk[j]) << 32) | vv[j];
                }
            }
            __syncthreads();
            dbase[t] += in_tile;
        }
        if (pass < 3 && !xcd_barrier(bar, gen, nblk, err, kErrSortBarrier)) return;
        uint32_t *tk = sk; sk = dk; dk = tk;
        uint32_t *tv = sv; sv = dv; dv = tv;
    }
}

// ---------------------------------------------------------------------------
// 2. Karras build + escape links
// ---------------------------------------------------------------------------
// Internal nodes whose leaf ranges cross a refit chunk ("top" nodes) are
// listed for refit_top: top_slot[i] = the node's place in top_list, -1 for a
// node inside one chunk, -2 past kTopMax (refit_top then climbs in global
// memory).
constexpr uint32_t kTopMax = 1024;
__global__ void karras_build(const uint64_t *__restrict__ codes, int n, int32_t *__restrict__ left,
                             int32_t *__restrict__ right, int32_t *__restrict__ parent, int2 *__restrict__ range,
                             TNode *__restrict__ tn, uint32_t *__restrict__ cnt_top, int32_t *__restrict__ root,
                             int32_t *__restrict__ top_slot, uint32_t *__restrict__ top_list,
                             uint32_t *top_count) {
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n - 1) return;
    int lo, hi, g;
    karras_node(codes, n, i, lo, hi, g);
    const int l = (lo == g) ? (n - 1 + lo) : g;
    const int r = (g + 1 == hi) ? (n - 1 + g + 1) : (g + 1);
    left[i] = l;
    right[i] = r;
    parent[l] = i;
    parent[r] = i;
    range[i] = make_int2(lo, hi);
    cnt_top[i] = 0u;
    int32_t ts = -1;
    if (lo / static_cast<int>(kChunk) != hi / static_cast<int>(kChunk)) {
        const uint32_t slot = atomicAdd(top_count, 1u);
        ts = -2;
        if (slot < kTopMax) {
            top_list[slot] = static_cast<uint32_t>(i);
            ts = static_cast<int32_t>(slot);
        }
    }
    top_slot[i] = ts;
    tn[i].child = l;
    tn[l].escape = r;
    tn[r].escape = karras_escape(codes, n, hi);
    if (i == 0) {
        parent[0] = -1;
        tn[0].escape = kEnd;
        root[0] = 0;
    }
}

// ---------------------------------------------------------------------------
// 3. Leaf boxes + refit
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kChunk) void leaf_refit(dy_body_view bv, dy_shape_view sv, int n,
                                                     const uint32_t *__restrict__ sorted_idx,
                                                     const int32_t *__restrict__ parent,
                                                     const int32_t *__restrict__ left,
                                                     const int32_t *__restrict__ right,
                                                     const int2 *__restrict__ range, dy_node_box *nodebox,
                                                     TNode *tn, TPair *tp, uint32_t *cnt_top, uint32_t *entries,
                                                     uint32_t *counters /* [0] entries */) {
    __shared__ uint32_t lbox[2][kChunk][3];  // [0] leaves, [1] internal nodes
    __shared__ uint32_t lmax[2][kChunk];
    __shared__ uint32_t lcnt[kChunk];
    // topology of the internal nodes base .. base + kChunk - 1 (a node whose
    // range lies inside the chunk has its number inside it: a Karras node is
    // one end of its own range)
    __shared__ int32_t lpar[kChunk], lleft[kChunk], lright[kChunk];
    __shared__ int2 lrange[kChunk];
    const int tid = static_cast<int>(threadIdx.x);
    const int base = static_cast<int>(blockIdx.x * kChunk);
    const int k = base + tid;
    lcnt[tid] = 0u;
    if (k < n - 1) {
        lpar[tid] = parent[k];
        lleft[tid] = left[k];
        lright[tid] = right[k];
        lrange[tid] = range[k];
    }
    const int leaf_par = k < n ? parent[n - 1 + k] : -1;
    __syncthreads();
    if (k < n) {
        const uint32_t body = sorted_idx[k];
        const uint32_t sh = bv.shape[body];
        float ex, ey, ez;
        leaf_extent(sv.type[sh], sv.half_x[sh], sv.half_y[sh], sv.half_z[sh], bv.rot_w[body],
                    bv.rot_x[body], bv.rot_y[body], bv.rot_z[body], ex, ey, ez);
        const float px = bv.pos_x[body], py = bv.pos_y[body], pz = bv.pos_z[body];
        uint32_t p[3];
        p[0] = f32_to_f16(px - ex) | (static_cast<uint32_t>(f32_to_f16(py - ey)) << 16);
        p[1] = f32_to_f16(pz - ez) | (static_cast<uint32_t>(f32_to_f16(px + ex)) << 16);
        p[2] = f32_to_f16(py + ey) | (static_cast<uint32_t>(f32_to_f16(pz + ez)) << 16);
        const int leaf = n - 1 + k;
        store_node(nodebox, tn, leaf, p, body);
        tn[leaf].child = static_cast<int32_t>(body);
        lbox[0][tid][0] = p[0]; lbox[0][tid][1] = p[1]; lbox[0][tid][2] = p[2];
        lmax[0][tid] = body;
        int node = leaf;
        for (int guard = 0;; ++guard) {
            if (guard >= kMaxHeight) { atomicOr(&counters[2], 1u); break; }
            const int par = node == leaf ? leaf_par : lpar[node - base];
            if (par < 0) break;
            const bool inside = par >= base && par < base + static_cast<int>(kChunk) && par < n - 1;
            const int2 rg = inside ? lrange[par - base] : make_int2(-1, -1);
            if (!inside || rg.x < base || rg.y >= base + static_cast<int>(kChunk)) {
                const uint32_t e = atomicAdd(&counters[0], 1u);
                if (e < static_cast<uint32_t>(n)) entries[e] = static_cast<uint32_t>(node);
                else atomicOr(&counters[2], 2u);
                break;
            }
            const uint32_t old = __hip_atomic_fetch_add(&lcnt[par - base], 1u, __ATOMIC_ACQ_REL,
                                                        __HIP_MEMORY_SCOPE_WORKGROUP);
            if (old == 0u) break;
            const int c[2] = {lleft[par - base], lright[par - base]};
            uint32_t cb[2][3], cm[2];
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const bool is_leaf = c[s] >= n - 1;
                const int slot = is_leaf ? (c[s] - (n - 1) - base) : (c[s] - base);
                const int kind = is_leaf ? 0 : 1;
                cb[s][0] = lbox[kind][slot][0];
                cb[s][1] = lbox[kind][slot][1];
                cb[s][2] = lbox[kind][slot][2];
                cm[s] = lmax[kind][slot];
            }
            uint32_t m[3];
            merge_packed(cb[0], cb[1], m);
            const uint32_t mb = cm[0] > cm[1] ? cm[0] : cm[1];
            store_node(nodebox, tn, par, m, mb);
            store_pair(tp, par, cb[0], cm[0], cb[1], cm[1], c[0], c[1]);
            lbox[1][par - base][0] = m[0]; lbox[1][par - base][1] = m[1]; lbox[1][par - base][2] = m[2];
            lmax[1][par - base] = mb;
            node = par;
        }
    }
}

// The nodes whose ranges cross chunks, bottom-up from the chunk subtrees'
// roots that leaf_refit listed (one workgroup; the kernel boundary makes
// every chunk's nodes visible).  Children counters cnt_top at workgroup
// scope: the second child to arrive merges and climbs on.
__global__ __launch_bounds__(kChunk) void refit_top(int n, const int32_t *__restrict__ parent,
                                                    const int32_t *__restrict__ left,
                                                    const int32_t *__restrict__ right, dy_node_box *nodebox,
                                                    TNode *tn, TPair *tp, uint32_t *cnt_top,
                                                    const uint32_t *__restrict__ entries,
                                                    const int32_t *__restrict__ top_slot,
                                                    const uint32_t *__restrict__ top_list,
                                                    uint32_t *counters /* [0] entries, [1] top nodes */) {
    const int tid = static_cast<int>(threadIdx.x);
    const uint32_t n_top = counters[1];
    if (n_top <= kTopMax) {
        // All top nodes in LDS: each one's two child boxes (a child inside a
        // chunk: loaded from tn; a top child: filled in when it is merged),
        // its parent's slot and side, and a count of its top children.
        __shared__ uint32_t box[kTopMax][2][4];  // fp16 box words, max body
        __shared__ int32_t tnid[kTopMax], tpar[kTopMax], tchild[kTopMax][2];
        __shared__ uint32_t tneed[kTopMax], tcnt[kTopMax];
        for (uint32_t sl = static_cast<uint32_t>(tid); sl < n_top; sl += kChunk) {
            const int node = static_cast<int>(top_list[sl]);
            tnid[sl] = node;
            const int c[2] = {left[node], right[node]};
            const int p = parent[node];
            int32_t pp = -1;
            if (p >= 0) pp = (top_slot[p] << 1) | (left[p] == node ? 0 : 1);  // a top node's parent is a top node
            tpar[sl] = pp;
            uint32_t need = 0;
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                tchild[sl][side] = c[side];
                const bool top_child = c[side] < n - 1 && top_slot[c[side]] >= 0;
                if (top_child) {
                    ++need;
                } else {
                    box[sl][side][0] = tn[c[side]].b01;
                    box[sl][side][1] = tn[c[side]].b23;
                    box[sl][side][2] = tn[c[side]].b45;
                    box[sl][side][3] = tn[c[side]].maxbody;
                }
            }
            tneed[sl] = need;
            tcnt[sl] = 0u;
        }
        __syncthreads();
        for (uint32_t sl = static_cast<uint32_t>(tid); sl < n_top; sl += kChunk) {
            if (tneed[sl] != 0u) continue;
            int cur = static_cast<int>(sl);
            for (int guard = 0;; ++guard) {
                if (guard >= kMaxHeight) { atomicOr(&counters[2], 1u); break; }
                const uint32_t a[3] = {box[cur][0][0], box[cur][0][1], box[cur][0][2]};
                const uint32_t b[3] = {box[cur][1][0], box[cur][1][1], box[cur][1][2]};
                const uint32_t ml = box[cur][0][3], mr = box[cur][1][3];
                uint32_t m[3];
                merge_packed(a, b, m);
                const uint32_t mb = ml > mr ? ml : mr;
                const int node = tnid[cur];
                store_node(nodebox, tn, node, m, mb);
                store_pair(tp, node, a, ml, b, mr, tchild[cur][0], tchild[cur][1]);
                const int32_t pp = tpar[cur];
                if (pp < 0) break;
                const int ps = pp >> 1, side = pp & 1;
                box[ps][side][0] = m[0];
                box[ps][side][1] = m[1];
                box[ps][side][2] = m[2];
                box[ps][side][3] = mb;
                const uint32_t old = __hip_atomic_fetch_add(&tcnt[ps], 1u, __ATOMIC_ACQ_REL,
                                                            __HIP_MEMORY_SCOPE_WORKGROUP);
                if (old + 1u < tneed[ps]) break;
                cur = ps;
            }
        }
        __syncthreads();
        if (tid == 0) {
            counters[0] = 0u;
            counters[1] = 0u;
        }
        return;
    }
    uint32_t n_entries = counters[0];
    n_entries = n_entries < static_cast<uint32_t>(n) ? n_entries : static_cast<uint32_t>(n);
    for (uint32_t e = tid; e < n_entries; e += kChunk) {
        int node = static_cast<int>(entries[e]);
        for (int guard = 0;; ++guard) {
            if (guard >= kMaxHeight) { atomicOr(&counters[2], 1u); break; }
            const int par = parent[node];
            if (par < 0) break;
            const uint32_t old = __hip_atomic_fetch_add(&cnt_top[par], 1u, __ATOMIC_ACQ_REL,
                                                        __HIP_MEMORY_SCOPE_WORKGROUP);
            if (old == 0u) break;
            const int l = left[par], r = right[par];
            const uint32_t a[3] = {tn[l].b01, tn[l].b23, tn[l].b45};
            const uint32_t b[3] = {tn[r].b01, tn[r].b23, tn[r].b45};
            uint32_t m[3];
            merge_packed(a, b, m);
            const uint32_t ml = tn[l].maxbody, mr = tn[r].maxbody;
            store_node(nodebox, tn, par, m, ml > mr ? ml : mr);
            store_pair(tp, par, a, ml, b, mr, l, r);
            node = par;
        }
    }
    __syncthreads();
    if (tid == 0) {
        counters[0] = 0u;
        counters[1] = 0u;
    }
}

// ---------------------------------------------------------------------------
// 4-6. Traversal and ordered pair emission
// ---------------------------------------------------------------------------
// Insert x into the ascending register list s (keeps the kRow smallest);
// returns the element that no longer fits (all ones while the list has room).
__device__ __forceinline__ uint32_t row_insert(uint32_t (&s)[kRow], uint32_t x) {
#pragma unroll
    for (uint32_t j = 0; j < kRow; ++j) {
        const uint32_t lo = s[j] < x ? s[j] : x;
        x = s[j] < x ? x : s[j];
        s[j] = lo;
    }
    return x;
}

// Walk the tree for query box q; count partners with body index > thr and
// keep the kRow smallest of them in s (the rest go to ovf when given).
// Depth-first over internal nodes: a node is visited once its own box has
// passed the test at its parent, and its record tests both children, so each
// overlapping internal node is loaded once and no leaf is loaded.  The next
// node stays in a register; a per-lane stack in LDS (stk: this lane's column
// of a [kStack][kWalkThreads] array) holds the right children deferred while
// the left subtree is walked, so LDS is touched only where both children
// pass and where a subtree ends.  kStack entries hold any tree of at most
// kStack levels; a deeper one flags err.
constexpr uint32_t kWalkThreads = 64;
constexpr uint32_t kStack = 68;
template <uint32_t STRIDE = kWalkThreads>
__device__ __forceinline__ uint32_t walk(const TNode *__restrict__ tn, const TPair *__restrict__ tp, int n,
                                         const Box6 &q, uint32_t thr, uint32_t (&s)[kRow], uint32_t *stk,
                                         uint32_t *err, uint32_t *ovf = nullptr) {
    uint32_t count = 0;
#pragma unroll
    for (uint32_t j = 0; j < kRow; ++j) s[j] = 0xffffffffu;
    const int n_int = n - 1;
    int x = -1;
    {
        const uint4 r = *reinterpret_cast<const uint4 *>(&tn[0]);
        if (r.w > thr && overlaps(unpack_box(r.x, r.y, r.z), q)) x = 0;
    }
    uint32_t sp = 0;
    while (x >= 0) {
        if (x >= n_int) { atomicOr(err, 4u); break; }
        const uint4 w0 = *reinterpret_cast<const uint4 *>(&tp[x]);
        const uint4 w1 = *(reinterpret_cast<const uint4 *>(&tp[x]) + 1);
        const int2 ch = *reinterpret_cast<const int2 *>(&tp[x].left);
        const bool pl = w0.w > thr && overlaps(unpack_box(w0.x, w0.y, w0.z), q);
        const bool pr = w1.w > thr && overlaps(unpack_box(w1.x, w1.y, w1.z), q);
        const bool il = ch.x < n_int, ir = ch.y < n_int;
        if (pl && !il) {
            const uint32_t ev = row_insert(s, w0.w);
            if (ovf && ev != 0xffffffffu && count - kRow < kOvf) ovf[count - kRow] = ev;
            ++count;
        }
        if (pr && !ir) {
            const uint32_t ev = row_insert(s, w1.w);
            if (ovf && ev != 0xffffffffu && count - kRow < kOvf) ovf[count - kRow] = ev;
            ++count;
        }
        const bool gl = pl && il, gr = pr && ir;
        if (gl && gr) {  // defer the right subtree
            if (sp >= kStack) { atomicOr(err, 4u); break; }
            stk[(sp++) * STRIDE] = static_cast<uint32_t>(ch.y);
            x = ch.x;
        } else if (gl) {
            x = ch.x;
        } else if (gr) {
            x = ch.y;
        } else {
            x = sp > 0 ? static_cast<int>(stk[(--sp) * STRIDE]) : -1;
        }
    }
    return count;
}

__global__ __launch_bounds__(kWalkThreads) void traverse_count(const TNode *__restrict__ tn,
                                                               const TPair *__restrict__ tp, int n,
                                                               uint32_t *__restrict__ counts,
                                                               uint32_t *__restrict__ row, uint32_t *__restrict__ ovf,
                                                               uint32_t *__restrict__ leaf_of, uint32_t *err,
                                                               uint32_t xcd_map) {
    __shared__ uint32_t stk[kStack * kWalkThreads];
    // With xcd_map the launch's workgroups, which the hardware deals out to
    // the 8 XCDs round-robin, take contiguous runs of leaves per XCD: walks of
    // neighbouring leaves share subtrees, so each XCD's L2 holds one region.
    uint32_t g = blockIdx.x;
    if (xcd_map) {
        const uint32_t G = gridDim.x, per = G / kXcds, rem = G % kXcds, x = g % kXcds, slot = g / kXcds;
        g = x < rem ? x * (per + 1) + slot : rem * (per + 1) + (x - rem) * per + slot;
    }
    const int k = static_cast<int>(g * blockDim.x + threadIdx.x);
    if (k >= n) return;
    const TNode leaf = tn[n - 1 + k];
    const uint32_t a = static_cast<uint32_t>(leaf.child);
    const Box6 q = unpack_box(leaf.b01, leaf.b23, leaf.b45);
    uint32_t s[kRow];
    const uint32_t c = walk(tn, tp, n, q, a, s, stk + threadIdx.x, err, ovf + static_cast<size_t>(a) * kOvf);
    counts[a] = c;
    leaf_of[a] = static_cast<uint32_t>(k);
    const uint32_t m = c < kRow ? c : kRow;
    for (uint32_t j = 0; j < kRow; ++j)
        if (j < m) row[static_cast<size_t>(a) * kRow + j] = s[j];
}

// Team-cooperative traversal: kTeam lanes of a wavefront walk one query
// together.  The team keeps a shared stack of internal nodes whose boxes
// passed the test at their parent (LDS); each step pops up to kTeam of them,
// one per lane, loads their two-wide records and tests both children;
// passing internal children are pushed and passing leaves (bodies above the
// query's) appended to the team's partner list, at positions from team
// ballots (prefix counts), so a step costs one dependent record load for
// the whole team.  Any order finds the same partner set; at the end each
// partner's rank among them (a count of smaller ones) places it: ranks below
// kRow in the row, the rest in the overflow buffer, both in ascending order.
// A query whose stack (stack_cap entries) or partner list would outgrow the
// team's LDS is walked again by one lane with the per-lane walk (the team's
// stack area as its stack).
// A Karras tree's walk pushes each internal node at most once, so a team
// takes at most n - 1 steps; more flags err.
constexpr uint32_t kTeamBlock = 256;
constexpr uint32_t kTeamStack = 128;    // >= kStack: the fallback walk uses it
constexpr uint32_t kTeamPartners = kRow + kOvf;  // a longer row is completed by emit_tail's walks
static_assert(kTeamStack >= kStack && kTeamPartners <= kRow + kOvf, "team sizes");

template <uint32_t kTeam>
__global__ __launch_bounds__(kTeamBlock) void traverse_team(const TNode *__restrict__ tn,
                                                            const TPair *__restrict__ tp, int n,
                                                            uint32_t *__restrict__ counts,
                                                            uint32_t *__restrict__ row, uint32_t *__restrict__ ovf,
                                                            uint32_t *__restrict__ leaf_of, uint32_t *err,
                                                            uint32_t xcd_map, uint32_t stack_cap) {
    static_assert(64u % kTeam == 0u, "teams inside a wavefront");
    constexpr uint32_t kTeamsPerBlock = kTeamBlock / kTeam;
    __shared__ uint32_t stk_all[kTeamsPerBlock][kTeamStack];
    __shared__ uint32_t prt_all[kTeamsPerBlock][kTeamPartners];
    // contiguous runs of leaves per XCD, as traverse_count
    uint32_t g = blockIdx.x;
    if (xcd_map) {
        const uint32_t G = gridDim.x, per = G / kXcds, rem = G % kXcds, x = g % kXcds, slot = g / kXcds;
        g = x < rem ? x * (per + 1) + slot : rem * (per + 1) + (x - rem) * per + slot;
    }
    const uint32_t team = threadIdx.x / kTeam, tl = threadIdx.x % kTeam;
    const int k = static_cast<int>(g * kTeamsPerBlock + team);
    if (k >= n) return;  // whole teams
    uint32_t *stk = stk_all[team], *prt = prt_all[team];
    const TNode leaf = tn[n - 1 + k];
    const uint32_t a = static_cast<uint32_t>(leaf.child);
    const Box6 q = unpack_box(leaf.b01, leaf.b23, leaf.b45);
    const uint32_t lane = lane_id();
    const uint64_t tmask = ((1ull << kTeam) - 1ull) << (lane & ~(kTeam - 1u));
    const uint64_t below = tmask & ((1ull << lane) - 1ull);
    const int n_int = n - 1;
    uint32_t sp = 0, np = 0;
    bool full = false;
    {
        const uint4 r = *reinterpret_cast<const uint4 *>(&tn[0]);
        if (r.w > a && overlaps(unpack_box(r.x, r.y, r.z), q)) {
            if (tl == 0) stk[0] = 0u;
            sp = 1;
        }
    }
    uint32_t steps = 0;
    while (sp > 0) {
        if (++steps > static_cast<uint32_t>(n)) {
            if (tl == 0) atomicOr(err, 4u);
            return;
        }
        const uint32_t npop = sp < kTeam ? sp : kTeam;
        const int x = tl < npop ? static_cast<int>(stk[sp - 1u - tl]) : -1;
        sp -= npop;
        bool gl = false, gr = false, hl = false, hr = false;
        uint32_t cl = 0, cr = 0, vl = 0, vr = 0;
        if (x >= 0) {
            if (x >= n_int) {
                atomicOr(err, 4u);
            } else {
                const uint4 w0 = *reinterpret_cast<const uint4 *>(&tp[x]);
                const uint4 w1 = *(reinterpret_cast<const uint4 *>(&tp[x]) + 1);
                const int2 ch = *reinterpret_cast<const int2 *>(&tp[x].left);
                const bool pl = w0.w > a && overlaps(unpack_box(w0.x, w0.y, w0.z), q);
                const bool pr = w1.w > a && overlaps(unpack_box(w1.x, w1.y, w1.z), q);
                const bool il = ch.x < n_int, ir = ch.y < n_int;
                gl = pl && il;
                gr = pr && ir;
                hl = pl && !il;
                hr = pr && !ir;
                cl = static_cast<uint32_t>(ch.x);
                cr = static_cast<uint32_t>(ch.y);
                vl = w0.w;
                vr = w1.w;
            }
        }
        const uint64_t bgl = __ballot(gl) & tmask, bgr = __ballot(gr) & tmask;
        const uint64_t bhl = __ballot(hl) & tmask, bhr = __ballot(hr) & tmask;
        const uint32_t ngl = static_cast<uint32_t>(__popcll(bgl)), ngr = static_cast<uint32_t>(__popcll(bgr));
        const uint32_t nhl = static_cast<uint32_t>(__popcll(bhl)), nhr = static_cast<uint32_t>(__popcll(bhr));
        if (sp + ngl + ngr > stack_cap || np + nhl + nhr > kTeamPartners) {
            full = true;
            break;
        }
        if (gl) stk[sp + static_cast<uint32_t>(__popcll(bgl & below))] = cl;
        if (gr) stk[sp + ngl + static_cast<uint32_t>(__popcll(bgr & below))] = cr;
        if (hl) prt[np + static_cast<uint32_t>(__popcll(bhl & below))] = vl;
        if (hr) prt[np + nhl + static_cast<uint32_t>(__popcll(bhr & below))] = vr;
        sp += ngl + ngr;
        np += nhl + nhr;
    }
    if (full) {
        uint32_t c = 0;
        if (tl == 0) {
            uint32_t s[kRow];
            c = walk<1>(tn, tp, n, q, a, s, stk, err, ovf + static_cast<size_t>(a) * kOvf);
            counts[a] = c;
            leaf_of[a] = static_cast<uint32_t>(k);
            const uint32_t m = c < kRow ? c : kRow;
            for (uint32_t j = 0; j < kRow; ++j)
                if (j < m) row[static_cast<size_t>(a) * kRow + j] = s[j];
        }
        // A row of at most kRow + kOvf partners is emitted straight from the
        // overflow buffer, so the walk's overflow entries (in eviction
        // order) are put in ascending order, ranked in the team's LDS.
        c = __shfl(c, static_cast<int>(lane & ~(kTeam - 1u)), 64);
        if (c > kRow && c <= kRow + kOvf) {
            uint32_t *o = ovf + static_cast<size_t>(a) * kOvf;
            const uint32_t m2 = c - kRow;
            for (uint32_t i = tl; i < m2; i += kTeam) stk[i] = o[i];
            for (uint32_t i = tl; i < m2; i += kTeam) {
                const uint32_t v = stk[i];
                uint32_t r = 0;
                for (uint32_t j = 0; j < m2; ++j) r += stk[j] < v ? 1u : 0u;
                o[r] = v;
            }
        }
        return;
    }
    for (uint32_t j = tl; j < np; j += kTeam) {
        const uint32_t v = prt[j];
        uint32_t r = 0;
        for (uint32_t i = 0; i < np; ++i) r += prt[i] < v ? 1u : 0u;
        if (r < kRow) row[static_cast<size_t>(a) * kRow + r] = v;
        else ovf[static_cast<size_t>(a) * kOvf + (r - kRow)] = v;
    }
    if (tl == 0) {
        counts[a] = np;
        leaf_of[a] = static_cast<uint32_t>(k);
    }
}

// Entries kRow.. of a body's row (more than kRow partners): from the
// overflow buffer, kRow at a time in ascending order, or, past kRow + kOvf,
// by walks for the next kRow partners above the last one written.
__device__ __forceinline__ void emit_tail(const TNode *__restrict__ tn, const TPair *__restrict__ tp, int n,
                                          uint32_t a, uint32_t L, uint32_t off, const uint32_t *__restrict__ row,
                                          const uint32_t *__restrict__ ovf, const uint32_t *__restrict__ leaf_of,
                                          dy_contact_pair *__restrict__ pairs, uint32_t cap, uint32_t *stk,
                                          uint32_t *err) {
    uint32_t thr = row[static_cast<size_t>(a) * kRow + kRow - 1];
    uint32_t done = kRow;
    uint32_t s[kRow];
    if (L <= kRow + kOvf) {
        // Every further partner is in the overflow buffer (unordered, all
        // above the kept kRow): select them kRow at a time in order.
        const uint32_t *o = ovf + static_cast<size_t>(a) * kOvf;
        const uint32_t m2 = L - kRow;
        while (done < L && off + done < cap) {
#pragma unroll
            for (uint32_t j = 0; j < kRow; ++j) s[j] = 0xffffffffu;
            for (uint32_t j = 0; j < m2; ++j) {
                const uint32_t v = o[j];
                if (v > thr) row_insert(s, v);
            }
            const uint32_t c = (L - done) < kRow ? (L - done) : kRow;
            for (uint32_t j = 0; j < kRow; ++j)
                if (j < c && off + done + j < cap) pairs[off + done + j] = dy_contact_pair{a, s[j]};
            done += c;
            thr = s[kRow - 1];
        }
        return;
    }
    // Longer rows: fetch the next kRow partners above the last one written.
    const TNode leaf = tn[n - 1 + leaf_of[a]];
    const Box6 q = unpack_box(leaf.b01, leaf.b23, leaf.b45);
    while (done < L && off + done < cap) {
        walk(tn, tp, n, q, thr, s, stk, err);
        const uint32_t c = (L - done) < kRow ? (L - done) : kRow;
        for (uint32_t j = 0; j < kRow; ++j)
            if (j < c && off + done + j < cap) pairs[off + done + j] = dy_contact_pair{a, s[j]};
        done += c;
        thr = s[kRow - 1];
    }
}

// One thread per body: writes its row (a, b0 < b1 < ...).
__global__ __launch_bounds__(kWalkThreads) void emit_pairs(const TNode *__restrict__ tn,
                                                           const TPair *__restrict__ tp, int n,
                                                  const uint32_t *__restrict__ counts,
                                                  const uint32_t *__restrict__ offs, const uint32_t *__restrict__ row,
                                                  const uint32_t *__restrict__ ovf,
                                                  const uint32_t *__restrict__ leaf_of,
                                                  dy_contact_pair *__restrict__ pairs, uint32_t cap,
                                                  uint32_t *err) {
    __shared__ uint32_t stk[kStack * kWalkThreads];
    const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= static_cast<uint32_t>(n)) return;
    const uint32_t L = counts[a];
    if (L == 0) return;
    const uint32_t off = offs[a];
    if (off >= cap) return;
    const uint32_t m = L < kRow ? L : kRow;
    for (uint32_t j = 0; j < m; ++j) {
        if (off + j < cap) pairs[off + j] = dy_contact_pair{a, row[static_cast<size_t>(a) * kRow + j]};
    }
    if (L > kRow) emit_tail(tn, tp, n, a, L, off, row, ovf, leaf_of, pairs, cap, stk + threadIdx.x, err);
}

// One wavefront per 64 consecutive bodies, whose pairs are one contiguous
// range of the output: the 64 rows (contiguous) are staged in LDS with
// coalesced loads, then each lane writes every 64th pair of the range,
// finding its body by a binary search over the rows' offsets, so each store
// instruction writes 64 consecutive pairs.  traverse_team leaves the
// overflow entries of a row in ascending order, so entries kRow .. kRow +
// kOvf come straight from the overflow buffer; rows longer than that are
// completed by the body's lane (emit_tail's walks).
__global__ __launch_bounds__(kWalkThreads) void emit_coop(const TNode *__restrict__ tn,
                                                          const TPair *__restrict__ tp, int n,
                                                          const uint32_t *__restrict__ counts,
                                                          const uint32_t *__restrict__ offs,
                                                          const uint32_t *__restrict__ row,
                                                          const uint32_t *__restrict__ ovf,
                                                          const uint32_t *__restrict__ leaf_of,
                                                          dy_contact_pair *__restrict__ pairs, uint32_t cap,
                                                          uint32_t *err) {
    __shared__ uint32_t stk[kStack * kWalkThreads];
    __shared__ uint4 lrow[kWalkThreads * kRow / 4];
    __shared__ uint32_t loff[kWalkThreads + 1];
    const uint32_t lane = threadIdx.x;
    const uint32_t a0 = blockIdx.x * kWalkThreads;
    const uint32_t nb = static_cast<uint32_t>(n) - a0 < kWalkThreads ? static_cast<uint32_t>(n) - a0 : kWalkThreads;
    const uint32_t a = a0 + lane;
    const bool valid = lane < nb;
    const uint32_t L = valid ? counts[a] : 0u, off = valid ? offs[a] : 0u;
    const uint4 *src = reinterpret_cast<const uint4 *>(row + static_cast<size_t>(a0) * kRow);
    for (uint32_t t = lane; t < nb * (kRow / 4); t += kWalkThreads) lrow[t] = src[t];
    const uint32_t base = __shfl(off, 0, 64);
    const uint32_t total = __shfl(off + L, static_cast<int>(nb - 1u), 64) - base;
    loff[lane] = valid ? off - base : total;
    if (lane == 0) loff[kWalkThreads] = total;
    __syncthreads();
    const uint32_t *lr = reinterpret_cast<const uint32_t *>(lrow);
    for (uint32_t e = lane; e < total; e += kWalkThreads) {
        if (base + e >= cap) break;
        uint32_t lo = 0, hi = nb;  // loff[lo] <= e < loff[hi]
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (loff[mid] <= e) lo = mid;
            else hi = mid;
        }
        const uint32_t j = e - loff[lo];
        const uint32_t b = a0 + lo;
        if (j < kRow) pairs[base + e] = dy_contact_pair{b, lr[lo * kRow + j]};
        else if (loff[lo + 1] - loff[lo] <= kRow + kOvf)
            pairs[base + e] = dy_contact_pair{b, ovf[static_cast<size_t>(b) * kOvf + (j - kRow)]};
    }
    if (valid && L > kRow + kOvf && off < cap)
        emit_tail(tn, tp, n, a, L, off, row, ovf, leaf_of, pairs, cap, stk + threadIdx.x, err);
}

}  // namespace dy

using namespace dy;

struct dy_broadphase_s {
    uint32_t max_bodies = 0, max_pairs = 0;
    uint32_t *code0 = nullptr, *idx0 = nullptr, *code1 = nullptr, *idx1 = nullptr;
    uint64_t *morton64 = nullptr;
    int32_t *left = nullptr, *right = nullptr, *parent = nullptr, *root = nullptr;
    int2 *range = nullptr;
    dy_node_box *nodebox = nullptr;
    TNode *tn = nullptr;
    uint32_t *cnt_top = nullptr, *entries = nullptr, *counters = nullptr, *top_list = nullptr;
    int32_t *top_slot = nullptr;
    uint32_t *counts = nullptr, *offs = nullptr, *row = nullptr, *ovf = nullptr, *leaf_of = nullptr;
    uint32_t *scan_tmp = nullptr;
    uint32_t *sort_hist = nullptr;  // morton_sort_grid's digit counts, 256 per tile
    dy_contact_pair *pairs = nullptr;
    uint32_t *count = nullptr;  // = counters + 3
    uint32_t sort_blocks = 0;   // participants of morton_sort_grid per launch (0: single-workgroup sort)
    uint32_t sort_cap = 0;      // the most one XCD holds at once
    bool sort_fixed = false;    // DYPHUR_SORT_BLOCKS given: sort_blocks as set, not one per tile
    TPair *tp = nullptr;        // walk records of the internal nodes
    // traverse_count / emit_pairs run 64-thread workgroups (kWalkThreads):
    // the walks spread over all CUs (256-thread ones left half of them idle at
    // 32k bodies); the XCD map keeps each XCD on one region of the tree.
    // DYPHUR_TRAVERSE_XCD=0 turns the map off.
    uint32_t trav_xcd = 1;
    uint32_t trav_team = 16;  // lanes per query of the team traversal (4, 8, 16; 0: traverse_count)
    uint32_t emit_coop = 1;  // wavefront-wide pair emission (DYPHUR_EMIT_COOP=0: emit_pairs)
    uint32_t team_stack = kTeamStack;  // team stack entries before the per-lane fallback (DYPHUR_TEAM_STACK)
    // Kernel timing (dy_broadphase_set_timing): events between the launches
    // of one build, read back when the next build starts or on request.
    bool timing = false, ev_pending = false;
    hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    dy_broadphase_timing acc{};
};

namespace {
void bp_free(dy_broadphase_s *b) {
    void *ptrs[] = {b->code0, b->idx0, b->code1, b->idx1, b->morton64, b->left, b->right, b->parent,
                    b->root, b->range, b->nodebox, b->tn, b->tp, b->cnt_top, b->entries, b->counters,
                    b->counts, b->offs, b->row, b->ovf, b->leaf_of, b->scan_tmp, b->pairs,
                    b->sort_hist, b->top_slot, b->top_list};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (hipEvent_t &e : b->ev)
        if (e) {
            (void)hipEventDestroy(e);
            e = nullptr;
        }
}
// Adds the event intervals of the last timed build to the totals.
int bp_collect_timing(dy_broadphase_s *b) {
    if (!b->ev_pending) return DY_OK;
    DY_TRY(hipEventSynchronize(b->ev[5]));
    float ms[5];
    for (int k = 0; k < 5; ++k) DY_TRY(hipEventElapsedTime(&ms[k], b->ev[k], b->ev[k + 1]));
    b->acc.builds += 1;
    b->acc.sort_ms += ms[0];
    b->acc.karras_ms += ms[1];
    b->acc.refit_ms += ms[2];
    b->acc.traverse_ms += ms[3];
    b->acc.emit_ms += ms[4];
    b->ev_pending = false;
    return DY_OK;
}
template <typename T>
hipError_t dalloc(T *&p, size_t n) {
    return hipMalloc(reinterpret_cast<void **>(&p), (n ? n : 1) * sizeof(T));
}
}  // namespace

extern "C" {

int dy_broadphase_create(dy_stream_t s, uint32_t max_bodies, uint32_t max_pairs, dy_broadphase_t *out) {
    if (!out) return DY_ERR_INVALID_ARG;
    *out = nullptr;
    dy_broadphase_s *b = new dy_broadphase_s;
    b->max_bodies = max_bodies;
    b->max_pairs = max_pairs;
    const size_t nb = max_bodies > 1 ? max_bodies : 1;
    const size_t nodes = 2 * nb - 1;
    hipError_t e = hipSuccess;
#define DY_A(ptr, cnt) \
    if (e == hipSuccess) e = dalloc(ptr, cnt)
    DY_A(b->code0, nb); DY_A(b->idx0, nb); DY_A(b->code1, nb); DY_A(b->idx1, nb);
    DY_A(b->morton64, nb);
    DY_A(b->left, nb); DY_A(b->right, nb); DY_A(b->parent, nodes); DY_A(b->root, 1);
    DY_A(b->range, nb); DY_A(b->nodebox, nodes); DY_A(b->tn, nodes); DY_A(b->tp, nb);
    DY_A(b->cnt_top, nb); DY_A(b->entries, nb); DY_A(b->counters, 8);
    DY_A(b->top_slot, nb); DY_A(b->top_list, kTopMax);
    DY_A(b->sort_hist, 256 * ceil_div(static_cast<uint32_t>(nb), kSgTile));
    DY_A(b->counts, nb); DY_A(b->offs, nb); DY_A(b->row, nb * kRow); DY_A(b->ovf, nb * kOvf); DY_A(b->leaf_of, nb);
    DY_A(b->scan_tmp, scan_tmp_words(nb)); DY_A(b->pairs, max_pairs);
#undef DY_A
    if (e == hipSuccess) b->count = b->counters + 3;
    if (e == hipSuccess) e = hipMemsetAsync(b->counters, 0, 8 * sizeof(uint32_t), hs(s));
    if (e == hipSuccess) {
        // Participating workgroups of morton_sort_grid (on one XCD): by
        // default one per 1 024-key tile up to kSortMaxBlocks (DYPHUR_SORT_BLOCKS
        // fixes the count), and no more than one XCD holds at once for a
        // launch of eight times as many.
        int dev = 0, coop = 0, per_cu = 0, cus = 0;
        e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e == hipSuccess && coop)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(morton_sort_grid),
                                                            kSgThreads, 0);
        if (e == hipSuccess && coop && per_cu > 0) {
            int want = static_cast<int>(kSortMaxBlocks);
            if (const char *env = std::getenv("DYPHUR_SORT_BLOCKS")) {
                want = std::atoi(env);
                b->sort_fixed = true;
            }
            const int cap = per_cu * cus / 8;
            b->sort_cap = static_cast<uint32_t>(cap > 0 ? cap : 0);
            b->sort_blocks = static_cast<uint32_t>(want < 0 ? 0 : (want > cap ? cap : want));
        }
    }
    if (const char *env = std::getenv("DYPHUR_TRAVERSE_XCD")) b->trav_xcd = std::atoi(env) ? 1u : 0u;
    if (const char *env = std::getenv("DYPHUR_EMIT_COOP")) b->emit_coop = std::atoi(env) ? 1u : 0u;
    if (const char *env = std::getenv("DYPHUR_TEAM_STACK")) {
        const int v = std::atoi(env);
        b->team_stack = v >= 1 && v <= static_cast<int>(kTeamStack) ? static_cast<uint32_t>(v) : kTeamStack;
    }
    if (const char *env = std::getenv("DYPHUR_TRAVERSE_TEAM")) {
        const int t = std::atoi(env);
        b->trav_team = (t == 4 || t == 8 || t == 16) ? static_cast<uint32_t>(t) : 0u;
    }
    if (e == hipSuccess) e = hipMemsetAsync(b->root, 0, sizeof(int32_t), hs(s));
    if (e != hipSuccess) {
        bp_free(b);
        delete b;
        return static_cast<int>(e);
    }
    *out = b;
    return DY_OK;
}

int dy_broadphase_destroy(dy_broadphase_t bp) {
    if (!bp) return DY_OK;
    (void)hipDeviceSynchronize();
    bp_free(bp);
    delete bp;
    return DY_OK;
}

int dy_broadphase_build_and_query(dy_broadphase_t bp, dy_stream_t s, const dy_body_view *bodies,
                                  const dy_shape_view *shapes, const dy_aabb *bounds) {
    if (!bp || !bodies || !shapes || !bounds) return DY_ERR_INVALID_ARG;
    const uint32_t n = bodies->n;
    if (n > bp->max_bodies || n >= kMortonMax) return DY_ERR_INVALID_ARG;
    hipStream_t st = hs(s);
    if (bp->timing) DY_TRY_STATUS(bp_collect_timing(bp));
    DY_TRY_STATUS(clear_words(st, bp->counters + 2, 6));
    if (n < 2) return DY_OK;
    const bool tm = bp->timing;
    if (tm) DY_TRY(hipEventRecord(bp->ev[0], st));
    const int ni = static_cast<int>(n);
    const uint32_t n_tiles = ceil_div(n, kSgTile);
    bool sorted = false;
    if (bp->sort_blocks > 0 && n_tiles > 4) {
        // sort_blocks participants on one XCD out of a launch with as many on
        // every XCD (8 on gfx950)
        // one participant per tile (each sorts one tile per pass), fewer when
        // the count is fixed or capped
        const uint32_t per_xcd = bp->sort_fixed || n_tiles > bp->sort_blocks ? bp->sort_blocks : n_tiles;
        const uint32_t blocks = 8u * per_xcd;
        dy_body_view bv = *bodies;
        dy_aabb bd = *bounds;
        uint32_t nv = n, ntv = n_tiles;
        uint32_t *c0 = bp->code0, *i0 = bp->idx0, *c1 = bp->code1, *i1 = bp->idx1, *hist = bp->sort_hist,
                 *bar = bp->counters + 4, *reg = bp->counters + 6, *err = bp->counters + 2;
        uint64_t *m64 = bp->morton64;
        void *args[] = {&bv, &bd, &nv, &ntv, &c0, &i0, &c1, &i1, &hist, &m64, &bar, &reg, &err};
        sorted = try_coop_launch(reinterpret_cast<const void *>(morton_sort_grid), dim3(blocks), dim3(kSgThreads),
                                 args, st) == hipSuccess;
        if (!sorted) bp->sort_blocks = 0;  // refused: the single-workgroup sort from now on
    }
    if (!sorted) {
        morton_sort<<<1, kMsThreads, 0, st>>>(*bodies, *bounds, n, bp->code0, bp->idx0, bp->code1, bp->idx1,
                                              bp->morton64);
        DY_LAUNCH_CHECK();
    }
    if (tm) DY_TRY(hipEventRecord(bp->ev[1], st));
    karras_build<<<ceil_div(n - 1, 256), 256, 0, st>>>(bp->morton64, ni, bp->left, bp->right, bp->parent,
                                                       bp->range, bp->tn, bp->cnt_top, bp->root, bp->top_slot,
                                                       bp->top_list, bp->counters + 1);
    DY_LAUNCH_CHECK();
    if (tm) DY_TRY(hipEventRecord(bp->ev[2], st));
    leaf_refit<<<ceil_div(n, kChunk), kChunk, 0, st>>>(*bodies, *shapes, ni, bp->idx0, bp->parent, bp->left,
                                                       bp->right, bp->range, bp->nodebox, bp->tn, bp->tp, bp->cnt_top,
                                                       bp->entries, bp->counters);
    DY_LAUNCH_CHECK();
    refit_top<<<1, kChunk, 0, st>>>(ni, bp->parent, bp->left, bp->right, bp->nodebox, bp->tn, bp->tp, bp->cnt_top,
                                    bp->entries, bp->top_slot, bp->top_list, bp->counters);
    DY_LAUNCH_CHECK();
    if (tm) DY_TRY(hipEventRecord(bp->ev[3], st));
    if (bp->trav_team == 4) {
        traverse_team<4><<<ceil_div(n, kTeamBlock / 4), kTeamBlock, 0, st>>>(
            bp->tn, bp->tp, ni, bp->counts, bp->row, bp->ovf, bp->leaf_of, bp->counters + 2, bp->trav_xcd,
            bp->team_stack);
    } else if (bp->trav_team == 8) {
        traverse_team<8><<<ceil_div(n, kTeamBlock / 8), kTeamBlock, 0, st>>>(
            bp->tn, bp->tp, ni, bp->counts, bp->row, bp->ovf, bp->leaf_of, bp->counters + 2, bp->trav_xcd,
            bp->team_stack);
    } else if (bp->trav_team == 16) {
        traverse_team<16><<<ceil_div(n, kTeamBlock / 16), kTeamBlock, 0, st>>>(
            bp->tn, bp->tp, ni, bp->counts, bp->row, bp->ovf, bp->leaf_of, bp->counters + 2, bp->trav_xcd,
            bp->team_stack);
    } else {
        traverse_count<<<ceil_div(n, kWalkThreads), kWalkThreads, 0, st>>>(
            bp->tn, bp->tp, ni, bp->counts, bp->row, bp->ovf, bp->leaf_of, bp->counters + 2, bp->trav_xcd);
    }
    DY_LAUNCH_CHECK();
    if (tm) DY_TRY(hipEventRecord(bp->ev[4], st));
    DY_TRY_STATUS(scan_exclusive(st, bp->counts, bp->offs, n, bp->count, bp->scan_tmp));
    if (bp->emit_coop && bp->trav_team)  // needs traverse_team's ordered overflow entries
        emit_coop<<<ceil_div(n, kWalkThreads), kWalkThreads, 0, st>>>(bp->tn, bp->tp, ni, bp->counts, bp->offs,
                                                                       bp->row, bp->ovf, bp->leaf_of, bp->pairs,
                                                                       bp->max_pairs, bp->counters + 2);
    else
        emit_pairs<<<ceil_div(n, kWalkThreads), kWalkThreads, 0, st>>>(bp->tn, bp->tp, ni, bp->counts, bp->offs,
                                                                        bp->row, bp->ovf, bp->leaf_of, bp->pairs,
                                                                        bp->max_pairs, bp->counters + 2);
    DY_LAUNCH_CHECK();
    if (tm) {
        DY_TRY(hipEventRecord(bp->ev[5], st));
        bp->ev_pending = true;
    }
    return DY_OK;
}

int dy_broadphase_set_timing(dy_broadphase_t bp, int enable) {
    if (!bp) return DY_ERR_INVALID_ARG;
    if (enable && !bp->ev[0])
        for (hipEvent_t &e : bp->ev) DY_TRY(hipEventCreate(&e));
    DY_TRY_STATUS(bp_collect_timing(bp));  // a build timed before this call is not counted
    bp->timing = enable != 0;
    bp->acc = dy_broadphase_timing{};
    return DY_OK;
}

int dy_broadphase_get_timing(dy_broadphase_t bp, dy_broadphase_timing *out) {
    if (!bp || !out) return DY_ERR_INVALID_ARG;
    DY_TRY_STATUS(bp_collect_timing(bp));
    *out = bp->acc;
    return DY_OK;
}

int dy_broadphase_pairs(dy_broadphase_t bp, const dy_contact_pair **pairs, const uint32_t **d_count,
                        uint32_t *max_pairs) {
    if (!bp) return DY_ERR_INVALID_ARG;
    if (pairs) *pairs = bp->pairs;
    if (d_count) *d_count = bp->count;
    if (max_pairs) *max_pairs = bp->max_pairs;
    return DY_OK;
}

int dy_broadphase_sort_pairs(dy_broadphase_t bp, dy_stream_t, uint32_t) {
    return bp ? DY_OK : DY_ERR_INVALID_ARG;
}

int dy_broadphase_download_count(dy_broadphase_t bp, dy_stream_t s, uint32_t *out) {
    if (!bp || !out) return DY_ERR_INVALID_ARG;
    uint32_t h[2];
    DY_TRY(hipMemcpyAsync(h, bp->counters + 2, sizeof(h), hipMemcpyDeviceToHost, hs(s)));
    DY_TRY(hipStreamSynchronize(hs(s)));
    *out = h[1];
    return h[0] ? DY_ERR_INTERNAL : DY_OK;
}

int dy_broadphase_bvh_view(dy_broadphase_t bp, uint32_t n_active, dy_bvh_view *out) {
    if (!bp || !out) return DY_ERR_INVALID_ARG;
    out->left = bp->left;
    out->right = bp->right;
    out->parent = bp->parent;
    out->root = bp->root;
    out->sorted_idx = bp->idx0;
    out->aabb = bp->nodebox;
    out->n = n_active;
    return DY_OK;
}

// Morton keys of the last build (leaf order); for parity tests.  Async.
int dy_broadphase_morton_keys(dy_broadphase_t bp, const uint64_t **keys) {
    if (!bp || !keys) return DY_ERR_INVALID_ARG;
    *keys = bp->morton64;
    return DY_OK;
}

}  // extern "C"
