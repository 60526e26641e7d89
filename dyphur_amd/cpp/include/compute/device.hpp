#pragma once
// Device (compute/include/compute/device.hpp:8-45) over the HIP devices the
// gfx950 library sees.  This build has no CPU device: default_cpu() throws,
// so code written for the reference's CPU backend fails loudly instead of
// running somewhere else; preferred() is the first GPU.
#include "stream.hpp"

#include <string>

namespace dyphur {

class Device {
public:
    explicit Device(int index) : index_(index) {
        int n = 0;
        check(dy_device_count(&n), "Device: dy_device_count");
        if (index < 0 || index >= n) throw Error(DY_ERR_NO_DEVICE, "Device: no HIP device " + std::to_string(index));
        check(dy_device_get_info(index, &info_), "Device: dy_device_get_info");
    }

    static Device default_gpu() { return Device(0); }
    static Device default_cpu() {
        throw Error(DY_ERR_NO_DEVICE, "Device::default_cpu: this build has no CPU device, it runs on gfx950 "
                                      "GPUs (use Device::default_gpu() or Device::preferred())");
    }
    static Device preferred() { return default_gpu(); }

    std::string name() const { return info_.name; }
    bool is_gpu() const noexcept { return info_.is_gpu != 0; }
    bool is_cpu() const noexcept { return false; }
    size_t max_compute_units() const noexcept { return info_.compute_units; }
    size_t global_mem_bytes() const noexcept { return info_.global_mem_bytes; }
    int index() const noexcept { return index_; }

    Stream make_stream() const { return Stream(*this); }

private:
    int index_ = 0;
    dy_device_info info_{};
};

inline Stream::Stream(const Device &dev) : device_(dev.index()) {
    check(dy_set_device(dev.index()), "Stream: dy_set_device");
    dy_stream_t s = nullptr;
    check(dy_stream_create(&s), "Stream: dy_stream_create");
    h_ = std::shared_ptr<dy_stream_s>(s, [](dy_stream_t p) { (void)dy_stream_destroy(p); });
    q_ = Queue(s, device_);
}

inline Device Queue::get_device() const { return Device(device_); }

}  // namespace dyphur
