// Per-island tables of the fused Newton kernel that depend only on the joint graph.
// They are built here, on the host, once per island when a world or an arena world
// is set up, and every step loads them (one 320-byte record per island) instead of
// rebuilding them on the device.
//
//   neighbour lists  — bodies coupled to body i in A (i itself first, then the other
//                      ends of its joints in joint order); the block sparsity the
//                      reference's assembly walks (crates/newton/src/newton.rs:300-360)
//   assembly terms   — per block (bi, bj): the (joint, row end, column end) triples
//                      whose Jacobian products add into it, in joint order
//   incidence        — per body: (joint, end) pairs in joint order (wrench sums)
//   joint order      — joints grouped by kind (thread mapping of the Jacobian lanes)
//   slab masks       — per 16-row tile of A: the 4-column slabs that can be non-zero
//   tile lists       — the 16×16 output tiles each wave computes in the two
//                      Newton–Schulz products
#include "island.h"

#include <cstring>
#include <vector>

namespace exg {

int build_island_topology(int m, int md, int J, const int* ja, const int* jb, const int* kind, int waves,
                          int per_wave, IslandTopo& t) {
    std::memset(&t, 0, sizeof t);
    if (m < 1 || m > TOPO_MAXM || md < 0 || md > m || J < 0 || J > TOPO_MAXJ || waves < 1 || waves > TOPO_WAVES ||
        per_wave < 1 || per_wave > TOPO_PER_WAVE)
        return -1;
    for (int j = 0; j < J; ++j)
        if (ja[j] < 0 || ja[j] >= m || jb[j] < 0 || jb[j] >= m || kind[j] < 0 || kind[j] > 3) return -1;
    // neighbour lists: the blocks of A, between dynamic bodies only (a kinematic body
    // has no row or column; its rows are empty lists)
    int lists[TOPO_MAXM][TOPO_MAXM], cnt[TOPO_MAXM];
    for (int i = 0; i < m; ++i) {
        lists[i][0] = i;
        cnt[i] = i < md ? 1 : 0;
    }
    auto add = [&](int i, int k) {
        if (i >= md || k >= md) return;
        for (int q = 0; q < cnt[i]; ++q)
            if (lists[i][q] == k) return;
        lists[i][cnt[i]++] = k;
    };
    for (int j = 0; j < J; ++j) {
        add(ja[j], jb[j]);
        add(jb[j], ja[j]);
    }
    int o = 0;
    for (int i = 0; i < m; ++i) {
        t.nbr_start[i] = (uint8_t)o;
        for (int q = 0; q < cnt[i]; ++q) t.nbr[o++] = (uint8_t)lists[i][q];
    }
    t.nbr_start[m] = (uint8_t)o;
    // assembly terms per block, blocks row-major over (bi, bj)
    o = 0;
    for (int b = 0; b < m * m; ++b) {
        t.term_start[b] = (uint8_t)o;
        const int bi = b / m, bj = b % m;
        for (int j = 0; j < J; ++j) {
            const int ends[2] = {ja[j], jb[j]};
            for (int re = 0; re < 2; ++re)
                for (int ce = 0; ce < 2; ++ce)
                    if (bi < md && bj < md && ends[re] == bi && ends[ce] == bj)
                        t.terms[o++] = (uint8_t)((j << 2) | (re << 1) | ce);
        }
    }
    t.term_start[m * m] = (uint8_t)o;
    // incidence
    o = 0;
    for (int i = 0; i < m; ++i) {
        t.inc_start[i] = (uint8_t)o;
        for (int j = 0; j < J; ++j) {
            if (ja[j] == i) t.inc[o++] = (uint8_t)(j << 1);
            if (jb[j] == i) t.inc[o++] = (uint8_t)((j << 1) | 1);
        }
    }
    t.inc_start[m] = (uint8_t)o;
    // joints by kind
    o = 0;
    for (int k = 0; k < 4; ++k)
        for (int j = 0; j < J; ++j)
            if (kind[j] == k) t.jorder[o++] = (uint8_t)j;
    // slab masks from the neighbour lists: body bi's rows meet body bk's columns
    for (int bi = 0; bi < m; ++bi)
        for (int q = t.nbr_start[bi]; q < t.nbr_start[bi + 1]; ++q) {
            const int bk = t.nbr[q];
            for (int tt = (6 * bi) / 16; tt <= (6 * bi + 5) / 16; ++tt)
                for (int ks = (6 * bk) / 4; ks <= (6 * bk + 5) / 4; ++ks) t.kmask[tt] |= (uint16_t)(1u << ks);
        }
    // tile lists: round robin for the dense product; for the sparse one, largest
    // count of non-zero slabs first, each to the least-loaded wave (fewest tiles on
    // a tie, then the lowest wave); round robin again if a wave would overflow
    const int T = (6 * md + 15) / 16, NTT = T * T;
    int rr[TOPO_WAVES][TOPO_PER_WAVE], rrn[TOPO_WAVES] = {};
    for (int tt = 0; tt < NTT; ++tt) {
        const int w = tt % waves;
        if (rrn[w] >= per_wave) return -1;
        rr[w][rrn[w]++] = tt;
    }
    int cost[TOPO_MAXT * TOPO_MAXT], order[TOPO_MAXT * TOPO_MAXT], load[TOPO_WAVES] = {};
    int gl[TOPO_WAVES][TOPO_PER_WAVE], gln[TOPO_WAVES] = {};
    for (int tt = 0; tt < NTT; ++tt) {
        int c = 0;
        for (unsigned v = t.kmask[tt / T]; v; v &= v - 1) ++c;
        cost[tt] = c;
        order[tt] = tt;
    }
    for (int x = 1; x < NTT; ++x)  // stable insertion sort, largest first
        for (int y = x; y > 0 && cost[order[y]] > cost[order[y - 1]]; --y) {
            const int tmp = order[y];
            order[y] = order[y - 1];
            order[y - 1] = tmp;
        }
    bool ok = true;
    for (int q = 0; q < NTT && ok; ++q) {
        int best = 0;
        for (int w = 1; w < waves; ++w)
            if (load[w] < load[best] || (load[w] == load[best] && gln[w] < gln[best])) best = w;
        if (gln[best] >= per_wave) {
            ok = false;
            break;
        }
        gl[best][gln[best]++] = order[q];
        load[best] += cost[order[q]];
    }
    for (int w = 0; w < waves; ++w) {
        t.tcount[1][w] = (uint8_t)rrn[w];
        t.tcount[0][w] = (uint8_t)(ok ? gln[w] : rrn[w]);
        for (int u = 0; u < rrn[w]; ++u) t.tlist[1][w][u] = (uint8_t)((rr[w][u] / T) | ((rr[w][u] % T) << 4));
        for (int u = 0; u < t.tcount[0][w]; ++u) {
            const int tt = ok ? gl[w][u] : rr[w][u];
            t.tlist[0][w][u] = (uint8_t)((tt / T) | ((tt % T) << 4));
        }
    }
    return 0;
}

int build_large_topology(int m, int md, int J, const int* ja, const int* jb, LargeTopo& t) {
    std::memset(&t, 0, sizeof t);
    if (m < 1 || m > LARGE_MAXM || md < 0 || md > m || J < 0 || J > LARGE_MAXJ) return -1;
    for (int j = 0; j < J; ++j)
        if (ja[j] < 0 || ja[j] >= m || jb[j] < 0 || jb[j] >= m || ja[j] == jb[j]) return -1;
    // neighbour lists: body i first, then the other ends of its joints in joint order
    // (between dynamic bodies only: a kinematic body has no row or column in A)
    std::vector<std::vector<int>> lists(m);
    for (int i = 0; i < md; ++i) lists[i].push_back(i);
    auto add = [&](int i, int k) {
        if (i >= md || k >= md) return;
        for (int x : lists[i])
            if (x == k) return;
        lists[i].push_back(k);
    };
    for (int j = 0; j < J; ++j) {
        add(ja[j], jb[j]);
        add(jb[j], ja[j]);
    }
    int o = 0;
    for (int i = 0; i < m; ++i) {
        t.nbr_start[i] = (uint16_t)o;
        for (int k : lists[i]) {
            if (o >= LARGE_MAXNBR) return -1;
            t.nbr[o] = (uint8_t)k;
            t.nbr_row[o] = (uint8_t)i;
            ++o;
        }
    }
    t.nbr_start[m] = (uint16_t)o;
    // assembly terms of each non-zero block (bi, bj), in joint order
    int to = 0;
    for (int q = 0; q < o; ++q) {
        t.term_start[q] = (uint16_t)to;
        const int bi = t.nbr_row[q], bj = t.nbr[q];
        for (int j = 0; j < J; ++j) {
            const int ends[2] = {ja[j], jb[j]};
            for (int re = 0; re < 2; ++re)
                for (int ce = 0; ce < 2; ++ce)
                    if (ends[re] == bi && ends[ce] == bj) t.terms[to++] = (uint16_t)((j << 2) | (re << 1) | ce);
        }
    }
    t.term_start[o] = (uint16_t)to;
    // incidence
    o = 0;
    for (int i = 0; i < m; ++i) {
        t.inc_start[i] = (uint16_t)o;
        for (int j = 0; j < J; ++j) {
            if (ja[j] == i) t.inc[o++] = (uint16_t)(j << 1);
            if (jb[j] == i) t.inc[o++] = (uint16_t)((j << 1) | 1);
        }
    }
    t.inc_start[m] = (uint16_t)o;
    // slab masks: body bi's rows meet body bk's columns
    for (int bi = 0; bi < m; ++bi)
        for (int q = t.nbr_start[bi]; q < t.nbr_start[bi + 1]; ++q) {
            const int bk = t.nbr[q];
            for (int tt = (6 * bi) / 16; tt <= (6 * bi + 5) / 16; ++tt)
                for (int ks = (6 * bk) / 4; ks <= (6 * bk + 5) / 4; ++ks) t.kmask[tt][ks >> 5] |= 1u << (ks & 31);
        }
    return 0;
}

long long large_workspace_doubles(int m, int J) {
    const long long n16 = ((6LL * m + 15) / 16) * 16;
    return 4 * n16 * n16 + 288LL * J;
}

int island_topology_of(int m, int md, int J, const int* kind, const int* a, const int* b, int b0, IslandTopo& t) {
    const int cls = island_class(m, J);
    if (cls < 0) return -1;
    if (cls == ISLAND_CLASS_M1 || cls == ISLAND_CLASS_LARGE) {   // no fixed record (large: LargeTopo)
        std::memset(&t, 0, sizeof t);
        return 0;
    }
    int waves = 0, per_wave = 0;
    island_class_shape(cls, waves, per_wave);
    int la[TOPO_MAXJ], lb[TOPO_MAXJ];
    if (J > TOPO_MAXJ) return -1;
    for (int j = 0; j < J; ++j) {
        la[j] = a[j] - b0;
        lb[j] = b[j] - b0;
    }
    return build_island_topology(m, md, J, la, lb, kind, waves, per_wave, t);
}

}  // namespace exg
