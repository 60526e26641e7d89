#pragma once
// ContactView (core/include/core/contact.hpp:8-20).
#include <dyphur_hip.h>

#include <cstddef>
#include <cstdint>

namespace dyphur {

struct ContactView {
    uint32_t *body_a;
    uint32_t *body_b;
    float *pos_x;
    float *pos_y;
    float *pos_z;
    float *norm_x;  // unit normal from B toward A
    float *norm_y;
    float *norm_z;
    float *depth;
    uint32_t *n;  // device pointer: live contact count
    uint32_t capacity;
};
static_assert(sizeof(ContactView) == sizeof(dy_contact_view) &&
                  offsetof(ContactView, capacity) == offsetof(dy_contact_view, capacity),
              "ContactView layout");

inline const dy_contact_view *c_view(const ContactView &v) { return reinterpret_cast<const dy_contact_view *>(&v); }

}  // namespace dyphur
