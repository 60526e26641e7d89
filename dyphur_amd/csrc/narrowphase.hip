// narrowphase.hip — primitive contact generation for gfx950
// (core/src/narrowphase.cpp:920-1127, sphere/box shapes).
//
// Each in-scope pair yields at most one contact (box-box manifolds reduce to
// their centroid), so the reference's append + (pair, sub) key sort + gather
// reduces to: one pair per lane into fixed per-pair slots with a hit flag and
// per-workgroup hit counts (np_test), a scan of the workgroup counts
// (np_offsets), and an ordered compaction (np_compact).  Contacts come out in
// pair order by construction; no sort, no atomics on the output.
#include "dy_internal.hpp"
#include "pair_contact.hpp"

namespace dy {

constexpr uint32_t kNpThreads = 256;

struct PairSlots {
    float *px, *py, *pz, *nx, *ny, *nz, *depth;
    uint32_t *hit;
};

__global__ __launch_bounds__(kNpThreads) void np_test(const dy_contact_pair *__restrict__ pairs, uint32_t n_bound,
                                                      const uint32_t *__restrict__ d_n, dy_body_view bv,
                                                      dy_shape_view sv, PairSlots slots,
                                                      uint32_t *__restrict__ block_hits, uint32_t *__restrict__ err) {
    __shared__ uint32_t red[kNpThreads / 64];
    const uint32_t i = blockIdx.x * kNpThreads + threadIdx.x;
    const uint32_t n = d_n ? min(*d_n, n_bound) : n_bound;
    uint32_t hit = 0, unsup = 0;
    if (i < n) {
        const dy_contact_pair p = pairs[i];
        Hit h;
        if (pair_contact(bv, sv, p.a, p.b, h, unsup)) {
            hit = 1;
            slots.px[i] = h.px; slots.py[i] = h.py; slots.pz[i] = h.pz;
            slots.nx[i] = h.nx; slots.ny[i] = h.ny; slots.nz[i] = h.nz;
            slots.depth[i] = h.depth;
        }
    }
    if (i < n_bound) slots.hit[i] = hit;
    if (unsup) atomicOr(err, 1u);
    uint32_t c = hit;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
    if (lane_id() == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) block_hits[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// counts[0] = stored (clamped), counts[1] = true count.
__global__ void np_clamp(uint32_t *__restrict__ counts, uint32_t cap) {
    const uint32_t t = counts[1];
    counts[0] = t < cap ? t : cap;
}

__global__ __launch_bounds__(kNpThreads) void np_compact(const dy_contact_pair *__restrict__ pairs, uint32_t n_bound,
                                                         PairSlots slots, const uint32_t *__restrict__ block_offs,
                                                         dy_contact_view cv) {
    __shared__ uint32_t wsum[17];
    const uint32_t i = blockIdx.x * kNpThreads + threadIdx.x;
    const uint32_t hit = i < n_bound ? slots.hit[i] : 0u;
    uint32_t total;
    const uint32_t pos = block_offs[blockIdx.x] + block_exclusive_scan(hit, wsum, total);
    if (!hit || pos >= cv.capacity) return;
    const dy_contact_pair p = pairs[i];
    cv.body_a[pos] = p.a;
    cv.body_b[pos] = p.b;
    cv.pos_x[pos] = slots.px[i]; cv.pos_y[pos] = slots.py[i]; cv.pos_z[pos] = slots.pz[i];
    cv.norm_x[pos] = slots.nx[i]; cv.norm_y[pos] = slots.ny[i]; cv.norm_z[pos] = slots.nz[i];
    cv.depth[pos] = slots.depth[i];
}

}  // namespace dy

using namespace dy;

struct dy_narrowphase_s {
    uint32_t cap = 0;
    // store
    uint32_t *body_a = nullptr, *body_b = nullptr;
    float *px = nullptr, *py = nullptr, *pz = nullptr, *nx = nullptr, *ny = nullptr, *nz = nullptr, *depth = nullptr;
    uint32_t *counts = nullptr;  // [0] stored, [1] true, [2] unsupported flag
    // per-pair scratch (grows)
    DevBuf<float> s_px, s_py, s_pz, s_nx, s_ny, s_nz, s_depth;
    DevBuf<uint32_t> s_hit, block_hits, block_offs, scan_tmp;
};

namespace {
int np_ensure(dy_narrowphase_s *np, uint32_t n_pairs) {
    const uint32_t nb = ceil_div(n_pairs > 0 ? n_pairs : 1, kNpThreads);
    DY_TRY_STATUS(np->s_px.ensure(n_pairs)); DY_TRY_STATUS(np->s_py.ensure(n_pairs));
    DY_TRY_STATUS(np->s_pz.ensure(n_pairs)); DY_TRY_STATUS(np->s_nx.ensure(n_pairs));
    DY_TRY_STATUS(np->s_ny.ensure(n_pairs)); DY_TRY_STATUS(np->s_nz.ensure(n_pairs));
    DY_TRY_STATUS(np->s_depth.ensure(n_pairs)); DY_TRY_STATUS(np->s_hit.ensure(n_pairs));
    DY_TRY_STATUS(np->block_hits.ensure(nb)); DY_TRY_STATUS(np->block_offs.ensure(nb));
    DY_TRY_STATUS(np->scan_tmp.ensure(scan_tmp_words(nb)));
    return DY_OK;
}

int np_launch(dy_narrowphase_s *np, hipStream_t st, const dy_contact_pair *pairs, uint32_t n_bound,
              const uint32_t *d_n, const dy_body_view &bv, const dy_shape_view &sv) {
    DY_TRY(hipMemsetAsync(np->counts, 0, 3 * sizeof(uint32_t), st));
    if (n_bound == 0) return DY_OK;
    DY_TRY_STATUS(np_ensure(np, n_bound));
    const uint32_t nb = ceil_div(n_bound, kNpThreads);
    PairSlots slots{np->s_px.p, np->s_py.p, np->s_pz.p, np->s_nx.p, np->s_ny.p, np->s_nz.p, np->s_depth.p, np->s_hit.p};
    np_test<<<nb, kNpThreads, 0, st>>>(pairs, n_bound, d_n, bv, sv, slots, np->block_hits.p, np->counts + 2);
    DY_LAUNCH_CHECK();
    DY_TRY_STATUS(scan_exclusive(st, np->block_hits.p, np->block_offs.p, nb, np->counts + 1, np->scan_tmp.p));
    np_clamp<<<1, 1, 0, st>>>(np->counts, np->cap);
    DY_LAUNCH_CHECK();
    dy_contact_view cv{np->body_a, np->body_b, np->px, np->py, np->pz, np->nx, np->ny, np->nz, np->depth,
                       np->counts, np->cap};
    np_compact<<<nb, kNpThreads, 0, st>>>(pairs, n_bound, slots, np->block_offs.p, cv);
    DY_LAUNCH_CHECK();
    return DY_OK;
}

void np_free(dy_narrowphase_s *np) {
    void *ptrs[] = {np->body_a, np->body_b, np->px, np->py, np->pz, np->nx, np->ny, np->nz, np->depth, np->counts};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    np->s_px.release(); np->s_py.release(); np->s_pz.release(); np->s_nx.release();
    np->s_ny.release(); np->s_nz.release(); np->s_depth.release(); np->s_hit.release();
    np->block_hits.release(); np->block_offs.release(); np->scan_tmp.release();
}
}  // namespace

extern "C" {

int dy_narrowphase_create(dy_stream_t s, uint32_t max_contacts, dy_narrowphase_t *out) {
    if (!out) return DY_ERR_INVALID_ARG;
    *out = nullptr;
    dy_narrowphase_s *np = new dy_narrowphase_s;
    np->cap = max_contacts;
    const size_t c = max_contacts ? max_contacts : 1;
    hipError_t e = hipSuccess;
    uint32_t **u[] = {&np->body_a, &np->body_b};
    float **f[] = {&np->px, &np->py, &np->pz, &np->nx, &np->ny, &np->nz, &np->depth};
    for (auto p : u)
        if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(p), c * sizeof(uint32_t));
    for (auto p : f)
        if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(p), c * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&np->counts), 4 * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemsetAsync(np->counts, 0, 4 * sizeof(uint32_t), hs(s));
    if (e != hipSuccess) {
        np_free(np);
        delete np;
        return static_cast<int>(e);
    }
    *out = np;
    return DY_OK;
}

int dy_narrowphase_destroy(dy_narrowphase_t np) {
    if (!np) return DY_OK;
    (void)hipDeviceSynchronize();
    np_free(np);
    delete np;
    return DY_OK;
}

int dy_narrowphase_run(dy_narrowphase_t np, dy_stream_t s, const dy_contact_pair *d_pairs, uint32_t n_pairs,
                       const dy_body_view *bodies, const dy_shape_view *shapes) {
    if (!np || !bodies || !shapes || (n_pairs && !d_pairs)) return DY_ERR_INVALID_ARG;
    return np_launch(np, hs(s), d_pairs, n_pairs, nullptr, *bodies, *shapes);
}

int dy_narrowphase_run_device_count(dy_narrowphase_t np, dy_stream_t s, const dy_contact_pair *d_pairs,
                                    const uint32_t *d_n_pairs, uint32_t max_pairs, const dy_body_view *bodies,
                                    const dy_shape_view *shapes) {
    if (!np || !bodies || !shapes || !d_n_pairs || (max_pairs && !d_pairs)) return DY_ERR_INVALID_ARG;
    return np_launch(np, hs(s), d_pairs, max_pairs, d_n_pairs, *bodies, *shapes);
}

int dy_narrowphase_contacts(dy_narrowphase_t np, dy_contact_view *out) {
    if (!np || !out) return DY_ERR_INVALID_ARG;
    *out = dy_contact_view{np->body_a, np->body_b, np->px, np->py, np->pz, np->nx, np->ny, np->nz, np->depth,
                           np->counts, np->cap};
    return DY_OK;
}

int dy_narrowphase_download_count(dy_narrowphase_t np, dy_stream_t s, uint32_t *out) {
    if (!np || !out) return DY_ERR_INVALID_ARG;
    uint32_t h[3];
    DY_TRY(hipMemcpyAsync(h, np->counts, sizeof(h), hipMemcpyDeviceToHost, hs(s)));
    DY_TRY(hipStreamSynchronize(hs(s)));
    *out = h[0];
    return h[2] ? DY_ERR_UNSUPPORTED : DY_OK;
}

int dy_narrowphase_last_count(dy_narrowphase_t np, dy_stream_t s, uint32_t *out) {
    if (!np || !out) return DY_ERR_INVALID_ARG;
    uint32_t h[3];
    DY_TRY(hipMemcpyAsync(h, np->counts, sizeof(h), hipMemcpyDeviceToHost, hs(s)));
    DY_TRY(hipStreamSynchronize(hs(s)));
    *out = h[1];
    return h[2] ? DY_ERR_UNSUPPORTED : DY_OK;
}

}  // extern "C"
