#pragma once
// Buffer<T> (compute/include/compute/buffer.hpp:11-55): move-only owner of
// n device elements; upload is asynchronous on the stream, download blocks.
#include "stream.hpp"

#include <cstddef>

namespace dyphur {

template <typename T>
class Buffer {
public:
    Buffer() = default;
    Buffer(Stream &s, size_t n) : s_(s.handle()), n_(n) {
        if (n > 0) {
            void *p = nullptr;
            check(dy_malloc(&p, n * sizeof(T)), "Buffer: dy_malloc");
            data_ = static_cast<T *>(p);
        }
    }
    ~Buffer() {
        if (data_) (void)dy_free(data_);
    }
    Buffer(const Buffer &) = delete;
    Buffer &operator=(const Buffer &) = delete;
    Buffer(Buffer &&o) noexcept : s_(o.s_), n_(o.n_), data_(o.data_) {
        o.data_ = nullptr;
        o.n_ = 0;
    }
    Buffer &operator=(Buffer &&o) noexcept {
        if (this != &o) {
            if (data_) (void)dy_free(data_);
            s_ = o.s_;
            n_ = o.n_;
            data_ = o.data_;
            o.data_ = nullptr;
            o.n_ = 0;
        }
        return *this;
    }

    T *data() noexcept { return data_; }
    const T *data() const noexcept { return data_; }
    size_t size() const noexcept { return n_; }
    bool empty() const noexcept { return n_ == 0; }

    Event upload(const T *src, size_t count = 0) {
        check(dy_memcpy_async(data_, src, (count ? count : n_) * sizeof(T), s_), "Buffer::upload");
        return Event::record(s_);
    }
    void download(T *dst, size_t count = 0) const {
        check(dy_memcpy_d2h(dst, data_, (count ? count : n_) * sizeof(T), s_), "Buffer::download");
    }

private:
    dy_stream_t s_ = nullptr;
    size_t n_ = 0;
    T *data_ = nullptr;
};

}  // namespace dyphur
