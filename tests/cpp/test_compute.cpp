// compute shim through the C++ layer (compute/tests/smoke_test.cpp:12-101,
// compute/tests/determinism_test.cpp:31-93).
#include "mini_test.hpp"

#include <compute/compute.hpp>

#include <cstdint>
#include <algorithm>
#include <numeric>
#include <vector>

using namespace dyphur;

TEST_CASE("parallel_for fills a buffer", "[compute]") {
    auto s = Device::default_gpu().make_stream();
    const size_t n = 1000;
    Buffer<float> b(s, n);
    float *d = b.data();
    parallel_for(s, n, [=](size_t i) { d[i] = 2.f * static_cast<float>(i); }).wait();
    std::vector<float> h(n);
    b.download(h.data());
    for (size_t i = 0; i < n; ++i) REQUIRE(h[i] == 2.f * static_cast<float>(i));
}

TEST_CASE("atomic_add counts every work item", "[compute]") {
    auto s = Device::default_gpu().make_stream();
    Buffer<uint32_t> c(s, 1);
    const uint32_t zero = 0;
    c.upload(&zero).wait();
    uint32_t *p = c.data();
    parallel_for(s, 4096, [=](size_t) { atomic_add(p, 1u); });
    parallel_for(s, 4096, [=](size_t) { atomic_add_seq(p, 1u); });
    uint32_t h = 0;
    c.download(&h);
    REQUIRE(h == 8192u);
}

TEST_CASE("sort_by_key: reverse order, u64 keys with u32 values", "[compute][sort]") {
    auto s = Device::default_gpu().make_stream();
    const size_t n = 4096;
    std::vector<uint64_t> k(n);
    std::vector<uint32_t> v(n);
    for (size_t i = 0; i < n; ++i) {
        k[i] = (static_cast<uint64_t>(n - i) << 32) | 7u;
        v[i] = static_cast<uint32_t>(i);
    }
    Buffer<uint64_t> dk(s, n);
    Buffer<uint32_t> dv(s, n);
    dk.upload(k.data());
    dv.upload(v.data());
    sort_by_key(s, dk.data(), dv.data(), n);
    s.wait();
    dk.download(k.data());
    dv.download(v.data());
    for (size_t i = 0; i < n; ++i) {
        REQUIRE(k[i] == ((static_cast<uint64_t>(i + 1) << 32) | 7u));
        REQUIRE(v[i] == static_cast<uint32_t>(n - 1 - i));
    }
}

TEST_CASE("sort: u32 keys with duplicates", "[compute][sort]") {
    auto s = Device::default_gpu().make_stream();
    const size_t n = 3000;
    std::vector<uint32_t> k(n);
    for (size_t i = 0; i < n; ++i) k[i] = static_cast<uint32_t>((i * 2654435761u) % 97u);
    Buffer<uint32_t> dk(s, n);
    dk.upload(k.data());
    sort(s, dk.data(), n);
    s.wait();
    std::vector<uint32_t> got(n);
    dk.download(got.data());
    std::sort(k.begin(), k.end());
    REQUIRE(got == k);
}

TEST_CASE("reduce_sum: ones, and the same bits every run", "[compute][reduce]") {
    auto s = Device::default_gpu().make_stream();
    const size_t n = 100003;
    std::vector<float> h(n, 1.f);
    Buffer<float> d(s, n);
    d.upload(h.data());
    REQUIRE_THAT(reduce_sum(s, d.data(), n), WithinAbs(static_cast<double>(n), 1e-5 * n));
    for (size_t i = 0; i < n; ++i) h[i] = 1.f / static_cast<float>(1 + i % 113);
    d.upload(h.data());
    const float a = reduce_sum(s, d.data(), n);
    for (int r = 0; r < 9; ++r) REQUIRE(reduce_sum(s, d.data(), n) == a);
}
