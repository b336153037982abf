// The host worker pool behind parallelFor (pale_scene.h): every item runs
// once; a loop inside a loop's task runs on the calling thread; loops from
// several threads at once all complete; an exception from a task reaches the
// caller and leaves the pool usable; a forked child, which has none of the
// parent's workers, still runs its loops.
#include <sys/wait.h>
#include <unistd.h>

#include <atomic>
#include <cstdio>
#include <stdexcept>
#include <thread>
#include <vector>

#include "pale_scene.h"

using pale::parallelFor;

static int fail(const char* what) {
    std::printf("FAIL %s\n", what);
    return 1;
}

int main() {
    for (int it = 0; it < 2000; ++it) {
        std::vector<int> v(1000, 0);
        parallelFor(v.size(), [&](size_t k) { v[k] += (int)k; });
        for (size_t k = 0; k < v.size(); ++k)
            if (v[k] != (int)k) return fail("each item once");
    }
    std::atomic<long> nested{0};
    parallelFor(64, [&](size_t a) { parallelFor(64, [&](size_t b) { nested += (long)(a * 64 + b); }); });
    if (nested != 4096L * 4095L / 2) return fail("nested loops");
    std::atomic<int> bad{0};
    std::vector<std::thread> callers;
    for (int c = 0; c < 4; ++c)
        callers.emplace_back([&] {
            for (int it = 0; it < 500; ++it) {
                std::atomic<long> s{0};
                parallelFor(100, [&](size_t k) { s += (long)k; });
                if (s != 4950) ++bad;
            }
        });
    for (auto& t : callers) t.join();
    if (bad) return fail("concurrent callers");
    for (int it = 0; it < 50; ++it) {
        bool caught = false;
        try {
            parallelFor(200, [&](size_t k) {
                if (k == 77) throw std::runtime_error("task");
            });
        } catch (const std::runtime_error&) {
            caught = true;
        }
        if (!caught) return fail("exception reaches the caller");
        std::atomic<long> s{0};
        parallelFor(100, [&](size_t k) { s += (long)k; });
        if (s != 4950) return fail("pool usable after an exception");
    }
    const pid_t child = fork();
    if (child == 0) {
        std::atomic<long> s{0};
        parallelFor(100, [&](size_t k) { s += (long)k; });
        _exit(s == 4950 ? 0 : 3);
    }
    int status = 0;
    waitpid(child, &status, 0);
    if (!WIFEXITED(status) || WEXITSTATUS(status) != 0) return fail("forked child");
    std::printf("ok\n");
    return 0;
}
