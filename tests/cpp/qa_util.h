// Minimal check harness of the C++ tests (no test framework in the image).
#pragma once

#include <cstdio>
#include <cstdlib>
#include <functional>
#include <string>
#include <vector>

static int g_failures = 0;

#define QA_CHECK(cond)                                                              \
    do {                                                                            \
        if (!(cond)) {                                                              \
            std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
            g_failures++;                                                           \
        }                                                                           \
    } while (0)

#define QA_CHECK_EQ(a, b)                                                           \
    do {                                                                            \
        const auto va_ = (a);                                                       \
        const auto vb_ = (b);                                                       \
        if (!(va_ == vb_)) {                                                        \
            std::fprintf(stderr, "%s:%d: %s == %s failed (%lld vs %lld)\n", __FILE__, __LINE__, #a, #b, \
                         (long long)va_, (long long)vb_);                           \
            g_failures++;                                                           \
        }                                                                           \
    } while (0)

struct qa_case {
    const char* name;
    std::function<void()> fn;
};

inline int qa_run(const std::vector<qa_case>& cases)
{
    for (const auto& c : cases) {
        const int before = g_failures;
        std::printf("RUN  %s\n", c.name);
        std::fflush(stdout);
        try {
            c.fn();
        } catch (const std::exception& e) {
            std::fprintf(stderr, "%s: exception: %s\n", c.name, e.what());
            g_failures++;
        }
        std::printf("%s %s\n", g_failures == before ? "PASS" : "FAIL", c.name);
        std::fflush(stdout);
    }
    std::printf("%d failure(s)\n", g_failures);
    return g_failures ? 1 : 0;
}
