// Checks apgk::sort_like_libstdcxx against std::sort on (key, tag) pairs
// compared by key only: the full sequence, ties included, must be identical.
// Inputs: random keys over few and many distinct values, sorted, reversed,
// constant, organ-pipe and sawtooth ranges, and McIlroy's adversary ("A
// Killer Adversary for Quicksort", Software: Practice and Experience 29(4),
// 1999) recorded against std::sort itself, which drives the introsort into
// its heapsort branch.  Prints "ok <cases>" or the first mismatch.
#include <stdio.h>

#include <algorithm>
#include <random>
#include <vector>

#include "prior_sort.hpp"

using apgk::KeyW;

static bool same_as_std(const std::vector<float>& keys, const char* what) {
    std::vector<KeyW> a(keys.size()), b;
    for (size_t i = 0; i < keys.size(); ++i) a[i] = KeyW{keys[i], (float)i};
    b = a;
    std::sort(a.begin(), a.end(), [](const KeyW& x, const KeyW& y) { return x.k < y.k; });
    apgk::sort_like_libstdcxx(b.data(), (int)b.size());
    for (size_t i = 0; i < a.size(); ++i)
        if (a[i].k != b[i].k || a[i].w != b[i].w) {
            printf("mismatch (%s, n=%zu) at %zu: std (%g, %g) here (%g, %g)\n", what, keys.size(), i, a[i].k, a[i].w,
                   b[i].k, b[i].w);
            return false;
        }
    return true;
}

// McIlroy's adversary: values are fixed lazily ("gas" until compared against
// another gas value), so std::sort meets its worst case; the frozen values
// are then a concrete input.
static std::vector<float> adversary(int n) {
    std::vector<int> val(n, n);
    int nsolid = 0, candidate = 0;
    const int gas = n;
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    auto freeze = [&](int x) { val[x] = nsolid++; };
    std::sort(idx.begin(), idx.end(), [&](int x, int y) {
        if (val[x] == gas && val[y] == gas) {
            if (x == candidate) freeze(x);
            else freeze(y);
        }
        if (val[x] == gas) candidate = x;
        else if (val[y] == gas) candidate = y;
        return val[x] < val[y];
    });
    std::vector<float> keys(n);
    for (int i = 0; i < n; ++i) keys[i] = (float)val[i];
    return keys;
}

int main() {
    std::mt19937 rng(2024);
    int cases = 0;
    auto run = [&](const std::vector<float>& k, const char* what) {
        ++cases;
        return same_as_std(k, what);
    };
    for (int n = 0; n <= 300; ++n)
        for (int distinct : {1, 2, 3, 5, 17, 1000}) {
            for (int rep = 0; rep < 4; ++rep) {
                std::vector<float> k(n);
                for (auto& v : k) v = (float)(rng() % (unsigned)distinct);
                if (!run(k, "random")) return 1;
            }
        }
    for (int n : {17, 31, 33, 64, 100, 257, 1000, 4096}) {
        std::vector<float> up(n), down(n), pipe(n), saw(n);
        for (int i = 0; i < n; ++i) {
            up[i] = (float)i;
            down[i] = (float)(n - i);
            pipe[i] = (float)std::min(i, n - 1 - i);
            saw[i] = (float)(i % 7);
        }
        if (!run(up, "sorted") || !run(down, "reversed") || !run(pipe, "organ pipe") || !run(saw, "sawtooth"))
            return 1;
        if (!run(std::vector<float>(n, 3.f), "constant")) return 1;
    }
    for (int n : {17, 18, 40, 64, 65, 100, 200, 513, 2000})
        if (!run(adversary(n), "adversary")) return 1;
    printf("ok %d\n", cases);
    return 0;
}
