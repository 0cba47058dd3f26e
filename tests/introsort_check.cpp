// Test harness (CPU): the permutation of nli_introsort.h's introsort_pairs against
// std::sort with the reference's comparator (patch_non_local_medians.h:41-47), on
// arrays with many equal keys.  Built and loaded by tests/test_introsort.py.
#include "nli_introsort.h"

#include <algorithm>
#include <random>
#include <utility>
#include <vector>

namespace {
struct CompareByKey {
    bool operator()(const std::pair<float, float> &a, const std::pair<float, float> &b) const { return a.first < b.first; }
};
}  // namespace

extern "C" int introsort_mismatches(unsigned seed, int cases, int max_n)
{
    std::mt19937 rng(seed);
    int bad = 0;
    for (int c = 0; c < cases; c++) {
        const int n = (int)(rng() % (unsigned)(max_n + 1));
        const int shape = (int)(rng() % 6);
        const int distinct = 1 + (int)(rng() % 12);
        std::vector<std::pair<float, float>> ref(n);
        std::vector<float> k(n), v(n);
        for (int i = 0; i < n; i++) {
            float key;
            switch (shape) {
            case 0: key = (float)(rng() % distinct); break;                    // few distinct keys
            case 1: key = 0.5f; break;                                          // all equal
            case 2: key = (float)(i / (1 + (int)(rng() % 4))); break;           // ascending runs
            case 3: key = (float)((n - i) / 3); break;                          // descending with ties
            case 4: key = (float)(std::min(i, n - i) % (distinct + 1)); break;  // organ pipe
            default: key = (float)(rng() % 1000) / 7.0f; break;                 // mostly distinct
            }
            k[i] = key;
            v[i] = (float)i;
            ref[i] = std::make_pair(key, (float)i);
        }
        std::sort(ref.begin(), ref.end(), CompareByKey());
        nli::introsort_pairs(k.data(), v.data(), n);
        for (int i = 0; i < n; i++)
            if (k[i] != ref[i].first || v[i] != ref[i].second) {
                bad++;
                break;
            }
    }
    return bad;
}

// Inputs that drive std::sort's quicksort to its depth budget (and so into the heapsort
// branch), with ties: McIlroy's adversary (A Killer Adversary for Quicksort, 1999) answers
// std::sort's own comparisons, fixing values lazily ("gas" items become "solid" as late as
// possible); frozen values are halved so that pairs of items tie.  The permutation of
// introsort_pairs is then compared with std::sort's on those keys.
namespace {
struct Adversary {
    std::vector<int> val;
    int gas, nsolid = 0, candidate = -1;
    explicit Adversary(int n) : val(n, n), gas(n) {}
    void freeze(int x) { val[x] = nsolid++ / 2; }
    bool operator()(int x, int y)
    {
        if (val[x] == gas && val[y] == gas) {
            if (x == candidate) freeze(x);
            else freeze(y);
        }
        if (val[x] == gas) candidate = x;
        else if (val[y] == gas) candidate = y;
        return val[x] < val[y];
    }
};
}  // namespace

extern "C" int introsort_adversarial_mismatches(int max_n)
{
    int bad = 0;
    for (int n = 17; n <= max_n; n++) {
        Adversary adv(n);
        std::vector<int> idx(n);
        for (int i = 0; i < n; i++) idx[i] = i;
        std::sort(idx.begin(), idx.end(), [&](int a, int b) { return adv(a, b); });
        std::vector<std::pair<float, float>> ref(n);
        std::vector<float> k(n), v(n);
        for (int i = 0; i < n; i++) {
            const float key = (float)(adv.val[i] == adv.gas ? adv.nsolid / 2 + 1 : adv.val[i]);
            ref[i] = std::make_pair(key, (float)i);
            k[i] = key;
            v[i] = (float)i;
        }
        std::sort(ref.begin(), ref.end(), CompareByKey());
        nli::introsort_pairs(k.data(), v.data(), n);
        for (int i = 0; i < n; i++)
            if (k[i] != ref[i].first || v[i] != ref[i].second) {
                bad++;
                break;
            }
    }
    return bad;
}

// median_order_free: whenever it holds, the weighted median over std::sort's order equals
// the one over the stable order (equal colours in input order); counts the cases it holds
// and the mismatches among them.  Weights are positive floats of varied magnitude.
namespace {
float walk(const std::vector<std::pair<float, float>> &m, float tw)
{
    float total = tw, med = 0.0f;
    const float half = tw / 2;
    for (const auto &p : m) {
        total -= p.second;
        if (total < half) { med = p.first; break; }
    }
    return med;
}
}  // namespace

extern "C" int median_order_free_check(unsigned seed, int cases, int *held)
{
    std::mt19937 rng(seed);
    int bad = 0;
    *held = 0;
    for (int c = 0; c < cases; c++) {
        int n = 2 + (int)(rng() % 120);
        const int distinct = 1 + (int)(rng() % 6);
        std::vector<std::pair<float, float>> a(n);
        for (int i = 0; i < n; i++) {
            const float w = (rng() % 4 == 0) ? 1e-3f * (float)(1 + rng() % 1000) : (float)(1 + rng() % 1000) / 81000.0f;
            a[i] = std::make_pair((float)(rng() % distinct), w);
        }
        if (c % 2 == 1) {
            // on the boundary: colour 1 gets the same weights as colour 0, so the exact total
            // after colour 0 is half the exact total, and rounding decides the crossing
            n = 2 * (n / 2 + 1);
            a.resize(n);
            for (int i = 0; i < n / 2; i++) {
                const float w = (float)(1 + rng() % 100000) / 8100000.0f;
                a[i] = std::make_pair(0.0f, w);
                a[n / 2 + i] = std::make_pair(1.0f, w);
            }
            std::shuffle(a.begin(), a.end(), rng);
        }
        float tw = 0.0f;
        for (int i = 0; i < n; i++) tw += a[i].second;
        std::vector<std::pair<float, float>> st = a, so = a;
        std::stable_sort(st.begin(), st.end(), CompareByKey());
        std::sort(so.begin(), so.end(), CompareByKey());
        std::vector<float> k(n), w(n);
        for (int i = 0; i < n; i++) { k[i] = st[i].first; w[i] = st[i].second; }
        if (nli::median_order_free(k.data(), w.data(), n, tw)) {
            (*held)++;
            if (walk(st, tw) != walk(so, tw)) bad++;
        }
    }
    return bad;
}
