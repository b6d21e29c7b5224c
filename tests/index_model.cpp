// Host model of the unconditional loads of the list build (exclusion slots),
// the integrator (contribution slots, list slots) and the bonded terms (member
// positions, tables), over edge-case layouts: no exclusions, no bonded terms,
// one atom, one bond, items with 8, 9 and 12 exclusions, term counts on both
// sides of a block.  Every device array is modelled with the size the library
// allocates (upload_count of the host array), and every index comes from the
// same functions the kernels use (ibisco_amd/csrc/ibx_index.h), so a fallback
// that leaves an array reads out of bounds here; built with AddressSanitizer.
// With --plant the exclusion fallback is replaced by the unclamped index, and
// the run must fail.
#define __host__
#define __device__
#include "ibx_index.h"

#include <cstdio>
#include <cstring>
#include <random>
#include <set>
#include <vector>

using namespace ibx;

static bool g_plant = false;

template <class T>
struct Dev {   // a device array: upload_count(n) elements, read through a raw pointer
    std::vector<T> v;
    explicit Dev(const std::vector<T> &h) : v(h) {
        if (v.empty()) v.assign((size_t)upload_count(0), T{});
    }
    T rd(long i) const { return v.data()[i]; }
};

struct Layout {
    int n = 0;                                  // items
    std::vector<std::vector<int>> bonds;        // adjacency
    std::vector<int> excl_off, excl;            // exclusions (CSR)
    std::vector<int> mem, meta;                 // terms, 4 members each, padded to whole blocks of 256
    int nterms = 0, ntabs = 0;
    std::vector<int> inc_off;                   // contribution slots of each item (CSR)
    int ncontrib = 0;
    std::vector<int> inv_perm;                  // slot of each item in a group (-1: not a member)
    int nslot = 0;
};

static Layout make(int n, const std::vector<std::pair<int, int>> &edges, int ntabs, int extra_excl_item = -1,
                   int extra_excl = 0, bool half_group = false) {
    Layout L;
    L.n = n;
    L.ntabs = ntabs;
    L.bonds.assign(n, {});
    for (auto &e : edges) {
        L.bonds[e.first].push_back(e.second);
        L.bonds[e.second].push_back(e.first);
    }
    // exclusions: 1-2 and 1-3 neighbours, plus extra ones for one item
    L.excl_off.assign(n + 1, 0);
    std::vector<std::vector<int>> ex(n);
    for (int i = 0; i < n; ++i) {
        std::set<int> s;
        for (int j : L.bonds[i]) {
            s.insert(j);
            for (int k : L.bonds[j])
                if (k != i) s.insert(k);
        }
        ex[i].assign(s.begin(), s.end());
    }
    if (extra_excl_item >= 0)
        for (int k = 0; (int)ex[extra_excl_item].size() < extra_excl && k < n; ++k)
            if (k != extra_excl_item) ex[extra_excl_item].push_back(k);
    for (int i = 0; i < n; ++i) {
        L.excl_off[i + 1] = L.excl_off[i] + (int)ex[i].size();
        for (int j : ex[i]) L.excl.push_back(j);
    }
    // terms: bonds and angles, each member owning one contribution slot
    std::vector<std::vector<int>> owned(n);
    auto add = [&](std::vector<int> m, int kind) {
        const int t = L.nterms++;
        for (int r = 0; r < 4; ++r) {
            const int it = r < (int)m.size() ? m[r] : -1;
            L.mem.push_back(it);
            if (it >= 0) owned[it].push_back(t);
        }
        L.meta.push_back(kind | ((ntabs > 0 ? t % ntabs : 0) << 2));
    };
    for (int i = 0; i < n; ++i)
        for (int j : L.bonds[i])
            if (j > i) add({i, j}, 0);
    for (int j = 0; j < n; ++j)
        for (size_t a = 0; a < L.bonds[j].size(); ++a)
            for (size_t b = a + 1; b < L.bonds[j].size(); ++b) add({L.bonds[j][a], j, L.bonds[j][b]}, 1);
    const int pad = (int)((L.nterms > 0 ? L.nterms : 1) + 255) / 256 * 256;
    while ((int)L.meta.size() < pad) {
        L.meta.push_back(-1);
        for (int r = 0; r < 4; ++r) L.mem.push_back(-1);
    }
    L.inc_off.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) L.inc_off[i + 1] = L.inc_off[i] + (int)owned[i].size();
    L.ncontrib = L.inc_off[n];
    // a group of every item, or of every second one
    L.inv_perm.assign(n, -1);
    for (int i = 0; i < n; ++i)
        if (!half_group || i % 2 == 0) L.inv_perm[i] = L.nslot++;
    return L;
}

static long reads = 0;

static void model(const Layout &L) {
    const Dev<int> excl(L.excl), inv_perm(L.inv_perm), excl_off(L.excl_off);
    // list build (k_build_list): eight exclusion slots loaded whatever the item's count
    for (int I = 0; I < L.n; ++I) {
        const int ex0 = excl_off.rd(I), ex1 = excl_off.rd(I + 1);
        int xi[8];
        for (int e = 0; e < 8; ++e) {
            const int k = g_plant ? ex0 + e : excl_entry(ex0, e, ex1);
            xi[e] = excl.rd(k);
            ++reads;
        }
        for (int e = 0; e < 8; ++e) {
            xi[e] = inv_perm.rd(excl_item(ex0, e, ex1, xi[e]));
            ++reads;
        }
        for (int e = ex0 + 8; e < ex1; ++e) (void)inv_perm.rd(excl.rd(e));
    }
    // integrator (k_move): the item's contribution slots, twelve at a time; its list slot in each group
    std::vector<float> hc((size_t)L.ncontrib * 4, 1.f);
    const Dev<float> contrib(hc);
    std::vector<int> hpos((size_t)L.nslot * 4, 0);
    const Dev<int> pos4_ref(hpos);
    const Dev<int> inc_off(L.inc_off);
    for (int it = 0; it < L.n; ++it) {
        const int a = inc_off.rd(it), b = inc_off.rd(it + 1);
        for (int q = a; q < b; q += 12)
            for (int u = 0; u < 12; ++u) {
                (void)contrib.rd(4L * contrib_entry(q + u, b));
                ++reads;
            }
        (void)pos4_ref.rd(4L * member_slot(inv_perm.rd(it)));
        ++reads;
    }
    // bonded terms (k_terms): every thread of the padded grid loads its four members and its table
    std::vector<double> hs(3 * (size_t)L.n, 0.0);
    const Dev<double> s(hs);
    const Dev<int> mem(L.mem), meta(L.meta);
    std::vector<int> htab((size_t)L.ntabs, 0);
    const Dev<int> tabs(htab);
    const int nthreads = (int)L.meta.size();
    for (int t = 0; t < nthreads; ++t) {
        const int mt = t < L.nterms ? meta.rd(t) : -1;
        for (int r = 0; r < 4; ++r) {
            const int m = mem.rd(4L * t + r);
            for (int k = 0; k < 3; ++k) (void)s.rd((long)k * L.n + term_member(mt, m));
            reads += 3;
        }
        (void)tabs.rd(term_table(mt));
        ++reads;
    }
}

int main(int argc, char **argv) {
    g_plant = argc > 1 && !strcmp(argv[1], "--plant");
    std::vector<Layout> cases;
    cases.push_back(make(2, {}, 0));                        // two atoms, nothing bonded, no tables
    cases.push_back(make(1, {}, 0));                        // one atom
    cases.push_back(make(2, {{0, 1}}, 1));                  // one bond
    cases.push_back(make(3, {{0, 1}, {1, 2}}, 2, -1, 0, true));   // one angle, every second item in the group
    cases.push_back(make(20, {{0, 1}, {1, 2}}, 1, 5, 8));   // an item with exactly 8 exclusions
    cases.push_back(make(20, {{0, 1}, {1, 2}}, 1, 5, 9));   // ... with 9
    cases.push_back(make(20, {{0, 1}, {1, 2}}, 1, 7, 12));  // ... with 12
    {
        std::vector<std::pair<int, int>> e;   // a chain of 129 bonds: 256 terms with its 127 angles, a full block
        for (int i = 0; i < 129; ++i) e.push_back({i, i + 1});
        cases.push_back(make(130, e, 3));
        e.push_back({130, 131});              // one bond and one angle more: the next block
        cases.push_back(make(132, e, 3));
    }
    {
        std::mt19937 rng(7);                  // random branched chains
        std::vector<std::pair<int, int>> e;
        for (int i = 1; i < 400; ++i) e.push_back({(int)(rng() % i), i});
        cases.push_back(make(400, e, 5, 0, 14, true));
    }
    for (const Layout &L : cases) model(L);
    printf("index model ok: %zu layouts, %ld unconditional loads in bounds\n", cases.size(), reads);
    return 0;
}
