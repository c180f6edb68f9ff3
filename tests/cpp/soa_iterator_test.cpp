// CPU test of the HostSideSOA proxy iteration (host_side_soa.h): the
// reference's own algorithms over SOAIterator (std::remove_if + std::sort, as
// in hssoa/host_side_soa.cpp:50-61, and std::lower_bound) must move whole
// points between the 16 SoA lines and agree with HostSideSOA's stable
// RemoveDisabledAndSort and FindFirstPointAtGridXIndex.
#include <algorithm>
#include <cstdio>
#include <random>
#include <set>
#include <tuple>
#include <vector>

#include "host_side_soa.h"

static int failures = 0;
#define CHECK(c)                                                           \
    do {                                                                   \
        if (!(c)) { std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c); failures++; } \
    } while (0)

// Wraps caller-owned memory (the destructor would free a pinned buffer).
struct Soa {
    std::vector<double> mem;
    HostSideSOA S;
    Soa(unsigned cap, unsigned n) : mem((size_t)cap * ProxyPoint::nArrays)
    {
        S.host_buffer = mem.data();
        S.capacity = cap;
        S.size = n;
    }
    ~Soa() { S.host_buffer = nullptr; }
};

// A point as the tuple of its 16 values (bit patterns of the utility word included).
static std::vector<double> point(HostSideSOA &S, unsigned i)
{
    std::vector<double> v;
    for (unsigned k = 0; k < ProxyPoint::nArrays; k++) v.push_back(S.getPointerToLine((int)k)[i]);
    return v;
}

int main()
{
    const unsigned cap = 1200, n = 1000, GridY = 64;
    const double h = 0.25, hinv = 1.0 / h;
    Soa a(cap, n), b(cap, n);
    std::mt19937_64 rng(0x1CE50A);
    std::uniform_real_distribution<double> ux(0.0, 40.0), uy(0.0, 15.0), u(0.0, 1.0);
    for (unsigned i = 0; i < n; i++) {
        int64_t ud = (int64_t)(i % 1000);
        if (i % 7 == 3) ud |= ICEMPM_STATUS_DISABLED;
        if (i % 11 == 5) ud |= ICEMPM_STATUS_CRUSHED;
        // coarse positions: many points share a cell (ties)
        const double x = std::floor(ux(rng) * 2) / 2, y = std::floor(uy(rng) * 2) / 2;
        for (HostSideSOA *S : {&a.S, &b.S}) {
            S->setUtility(i, ud);
            S->getPointerToPosX()[i] = x;
            S->getPointerToPosY()[i] = y;
            for (unsigned k = 1; k < ProxyPoint::nArrays; k++)
                if (k != icy::SimParams::posx && k != icy::SimParams::posx + 1) S->getPointerToLine((int)k)[i] = i + 0.001 * k;
        }
    }
    // proxy access
    SOAIterator it = a.S.begin() + 17;
    CHECK(it->getPos().x() == a.S.getPointerToPosX()[17] && (*it).getGrain() == 17);
    CHECK(it->getDisabledStatus() == ((17 % 7) == 3) && it->getCrushedStatus() == ((17 % 11) == 5));
    CHECK(a.S.end() - a.S.begin() == (int)n);
    CHECK(it->getCellIndex(hinv, GridY) == a.S.getCellIndex(17, hinv, GridY));
    {   // a value copy does not write through; assigning to a reference does
        ProxyPoint copy = *it;
        copy.setValue(icy::SimParams::Fe00, 42.0);
        CHECK(a.S.getPointerToLine((int)icy::SimParams::Fe00)[17] != 42.0);
        *(a.S.begin() + 18) = copy;
        CHECK(a.S.getPointerToLine((int)icy::SimParams::Fe00)[18] == 42.0);
        *(a.S.begin() + 18) = *(b.S.begin() + 18);            // restore from the twin
        CHECK(point(a.S, 18) == point(b.S, 18));
    }
    auto dims = a.S.getBlockDimensions();
    CHECK(dims.first.x() == *std::min_element(a.S.getPointerToPosX(), a.S.getPointerToPosX() + n));
    CHECK(dims.second.y() == *std::max_element(a.S.getPointerToPosY(), a.S.getPointerToPosY() + n));

    // the reference's algorithm through the iterator
    std::multiset<std::vector<double>> kept;
    for (unsigned i = 0; i < n; i++)
        if (!a.S.getDisabledStatus(i)) kept.insert(point(a.S, i));
    SOAIterator r = std::remove_if(a.S.begin(), a.S.end(), [](ProxyPoint &p) { return p.getDisabledStatus(); });
    a.S.size = r.m_point.pos;
    std::sort(a.S.begin(), a.S.end(),
              [&](ProxyPoint &p1, ProxyPoint &p2) { return p1.getCellIndex(hinv, GridY) < p2.getCellIndex(hinv, GridY); });
    b.S.RemoveDisabledAndSort(hinv, GridY);
    CHECK(a.S.size == b.S.size && a.S.size == kept.size());
    std::multiset<std::vector<double>> after;
    for (unsigned i = 0; i < a.S.size; i++) {
        after.insert(point(a.S, i));
        // same key sequence (the orders differ only among ties)
        CHECK(a.S.getCellIndex(i, hinv, GridY) == b.S.getCellIndex(i, hinv, GridY));
    }
    CHECK(after == kept);
    // lower bound on the x index
    for (int gx : {-3, 0, 1, 17, 80, 159, 160, 500}) {
        const unsigned f = a.S.FindFirstPointAtGridXIndex(gx, hinv);
        SOAIterator lb = std::lower_bound(a.S.begin(), a.S.end(), gx,
                                          [&](const ProxyPoint &p, int v) { return p.getXIndex(hinv) < v; });
        CHECK(f == lb.m_point.pos);
        unsigned brute = 0;
        while (brute < a.S.size && (int)(a.S.getPointerToPosX()[brute] * hinv + 0.5) < gx) brute++;
        CHECK(f == brute);
    }
    std::printf("%s: %d failures\n", failures ? "FAIL" : "ok", failures);
    return failures ? 1 : 0;
}
