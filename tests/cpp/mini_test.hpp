#pragma once
// A minimal Catch2-style harness (TEST_CASE / REQUIRE / REQUIRE_THAT with
// WithinAbs / REQUIRE_THROWS_AS) so the C++ API tests read like the
// reference's.  main() runs every case or, with --list, prints their names.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <functional>
#include <string>
#include <vector>

namespace mini {

struct Case {
    const char *name;
    const char *tags;
    std::function<void()> fn;
};
inline std::vector<Case> &registry() {
    static std::vector<Case> r;
    return r;
}
struct Reg {
    Reg(const char *n, const char *t, std::function<void()> f) { registry().push_back({n, t, std::move(f)}); }
};
struct Failure : std::exception {
    std::string msg;
    explicit Failure(std::string m) : msg(std::move(m)) {}
    const char *what() const noexcept override { return msg.c_str(); }
};
struct WithinAbs {
    double target, eps;
    bool match(double v) const { return std::fabs(v - target) <= eps; }
};
inline void fail(const char *file, int line, const std::string &expr) {
    throw Failure(std::string(file) + ":" + std::to_string(line) + ": REQUIRE(" + expr + ") failed");
}

inline int run_all(int argc, char **argv) {
    const bool list = argc > 1 && std::strcmp(argv[1], "--list") == 0;
    const char *only = (argc > 2 && std::strcmp(argv[1], "--only") == 0) ? argv[2] : nullptr;
    int failed = 0, ran = 0;
    for (auto &c : registry()) {
        if (list) {
            std::printf("%s\n", c.name);
            continue;
        }
        if (only && std::strstr(c.name, only) == nullptr) continue;
        ++ran;
        std::printf("[ RUN  ] %s\n", c.name);
        std::fflush(stdout);
        try {
            c.fn();
            std::printf("[  OK  ] %s\n", c.name);
        } catch (const std::exception &e) {
            ++failed;
            std::printf("[ FAIL ] %s: %s\n", c.name, e.what());
        }
        std::fflush(stdout);
    }
    if (!list) std::printf("%d of %d test cases passed\n", ran - failed, ran);
    return failed ? 1 : 0;
}

}  // namespace mini

#define MINI_CAT2(a, b) a##b
#define MINI_CAT(a, b) MINI_CAT2(a, b)
#define TEST_CASE(name, tags)                                                                   \
    static void MINI_CAT(mini_case_, __LINE__)();                                               \
    static mini::Reg MINI_CAT(mini_reg_, __LINE__)(name, tags, MINI_CAT(mini_case_, __LINE__)); \
    static void MINI_CAT(mini_case_, __LINE__)()
#define REQUIRE(e) \
    do { if (!(e)) mini::fail(__FILE__, __LINE__, #e); } while (0)
#define REQUIRE_FALSE(e) \
    do { if ((e)) mini::fail(__FILE__, __LINE__, "!(" #e ")"); } while (0)
#define REQUIRE_THAT(v, m)                                                                     \
    do {                                                                                       \
        const double mini_v_ = static_cast<double>(v);                                         \
        if (!(m).match(mini_v_))                                                               \
            mini::fail(__FILE__, __LINE__, std::string(#v " = ") + std::to_string(mini_v_) +   \
                                               " within " + std::to_string((m).eps) + " of " + \
                                               std::to_string((m).target));                    \
    } while (0)
#define REQUIRE_THROWS_AS(e, T)                                          \
    do {                                                                 \
        bool mini_thrown_ = false;                                       \
        try { (void)(e); } catch (const T &) { mini_thrown_ = true; }    \
        if (!mini_thrown_) mini::fail(__FILE__, __LINE__, "throws " #T); \
    } while (0)
using mini::WithinAbs;
