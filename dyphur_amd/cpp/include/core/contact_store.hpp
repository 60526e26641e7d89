#pragma once
// ContactStore (core/include/core/contact_store.hpp, core/src/contact_store.cpp:15-34).
#include "contact.hpp"

#include <compute/buffer.hpp>
#include <compute/stream.hpp>

namespace dyphur {

class ContactStore {
public:
    ContactStore() = default;
    ContactStore(Stream &s, uint32_t capacity)
        : cap_(capacity), d_u_(s, static_cast<size_t>(capacity) * 2 + 1), d_f_(s, static_cast<size_t>(capacity) * 7) {
        reset(s);
    }
    void reset(Stream &s) { check(dy_memset_async(count_ptr(), 0, sizeof(uint32_t), s.handle()), "ContactStore::reset"); }
    ContactView view() noexcept {
        const size_t c = cap_;
        float *f = d_f_.data();
        uint32_t *u = d_u_.data();
        return ContactView{u, u + c, f, f + c, f + 2 * c, f + 3 * c, f + 4 * c, f + 5 * c, f + 6 * c, u + 2 * c, cap_};
    }
    uint32_t capacity() const noexcept { return cap_; }
    uint32_t download_count(Stream &s) const {
        uint32_t n = 0;
        check(dy_memcpy_d2h(&n, d_u_.data() + 2 * static_cast<size_t>(cap_), sizeof n, s.handle()),
              "ContactStore::download_count");
        return n;
    }

private:
    uint32_t *count_ptr() noexcept { return d_u_.data() + 2 * static_cast<size_t>(cap_); }
    uint32_t cap_ = 0;
    Buffer<uint32_t> d_u_;  // body_a, body_b, count
    Buffer<float> d_f_;     // pos xyz, norm xyz, depth
};

}  // namespace dyphur
