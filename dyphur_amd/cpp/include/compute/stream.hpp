#pragma once
// Event / Queue / Stream (compute/include/compute/stream.hpp:6-30) over a HIP
// stream.  A Stream is in order: successive submissions run in submission
// order.  Copies of a Stream share one HIP stream, as copies of a sycl::queue
// share one queue.  queue() returns a Queue offering the memcpy / memset /
// wait / get_device calls that callers make on the reference's sycl::queue.
#include "status.hpp"

#include <cstddef>
#include <memory>

namespace dyphur {

class Device;

// Completion of one submission; wait() blocks until it has run and throws if
// the stream reported an error.
class Event {
public:
    Event() = default;
    static Event record(dy_stream_t s) {
        Event e;
        dy_event_t h = nullptr;
        check(dy_event_create(&h), "Event: dy_event_create");
        e.h_ = std::shared_ptr<dy_event_s>(h, [](dy_event_t p) { (void)dy_event_destroy(p); });
        check(dy_event_record(h, s), "Event: dy_event_record");
        return e;
    }
    void wait() {
        if (h_) check(dy_event_synchronize(h_.get()), "Event::wait");
    }
    dy_event_t handle() const noexcept { return h_.get(); }

private:
    std::shared_ptr<dy_event_s> h_;
};

class Queue {
public:
    Queue() = default;
    Queue(dy_stream_t s, int device) : s_(s), device_(device) {}
    // Asynchronous copy; the direction follows from the pointers.
    Event memcpy(void *dst, const void *src, size_t bytes) {
        check(dy_memcpy_async(dst, src, bytes, s_), "Queue::memcpy");
        return Event::record(s_);
    }
    Event memset(void *dst, int value, size_t bytes) {
        check(dy_memset_async(dst, value, bytes, s_), "Queue::memset");
        return Event::record(s_);
    }
    void wait() { check(dy_stream_synchronize(s_), "Queue::wait"); }
    void wait_and_throw() { wait(); }
    inline Device get_device() const;
    dy_stream_t handle() const noexcept { return s_; }

private:
    dy_stream_t s_ = nullptr;
    int device_ = 0;
};

class Stream {
public:
    inline explicit Stream(const Device &dev);

    Queue &queue() noexcept { return q_; }
    const Queue &queue() const noexcept { return q_; }
    void wait() { check(dy_stream_synchronize(handle()), "Stream::wait"); }
    dy_stream_t handle() const noexcept { return h_.get(); }
    int device() const noexcept { return device_; }

private:
    std::shared_ptr<dy_stream_s> h_;
    int device_ = 0;
    Queue q_;
};

}  // namespace dyphur

#include "device.hpp"
