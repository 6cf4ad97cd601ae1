// Host-side structure builder: cluster tree, near/far lists, coarsening,
// blocking, kNN sampling, interpolative decomposition and CDS packing.
//
// This is the host half of the evaluator (tree construction, sampling and FP64
// compression stay on the CPU).  Each stage restates the MatRox algorithm it
// cites; the output arrays keep the reference's names, index spaces and
// memory layout so that the evaluation phase sees exactly the structure the
// generated HGEMM would see (matroxTest-V1/main.cpp:64-107).
#include "../../../include/matrox_host.h"
#include "rng.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <unordered_set>
#include <utility>
#include <vector>

namespace mxh {

static thread_local std::string g_last_error;

struct Handle {
    int n = 0, d = 0;
    std::vector<double> X;
    mxh_config cfg{};

    int depth = 0;       // leaves sit at depth `depth` (root = depth 0)
    int numnodes = 0, nleaf = 0;
    bool is_hss = false;

    std::vector<std::vector<int>> pts;   // point ids of every node, in split order
    std::vector<double> mid, radius;     // bounding-box centre / radius per node
    std::vector<std::vector<int>> nearn, farn;
    std::vector<std::vector<int>> sflevel;               // bottom-up level sets
    std::vector<std::vector<int>> coarlevels;            // groups of sflevel indices
    std::vector<std::vector<std::vector<int>>> postw;    // post-order partitions
    std::vector<std::vector<std::vector<int>>> opostw;   // bin-packed partitions
    std::vector<std::pair<double, int>> NN;              // n*k, ascending per point
    int knn_k = 0;

    std::vector<std::vector<int>> skels;
    std::vector<std::vector<double>> proj;
    std::vector<int> projcol;
    std::vector<double> id_ratio;   // per node: |R_ss / R_00| of the first column the ID left out (0: none)
    std::vector<int> id_rows;       // per node: rows of the block the ID factored

    std::map<std::string, std::vector<int32_t>> i32;
    std::map<std::string, std::vector<uint64_t>> u64;
    std::map<std::string, std::vector<double>> f64;
    std::map<std::string, int64_t> scal;

    int leaf_first() const { return (1 << depth) - 1; }
    bool is_leaf(int v) const { return v >= leaf_first(); }
    int node_level(int v) const {  // depth of node v in the perfect tree
        int l = 0;
        while (v > 0) { v = (v - 1) / 2; ++l; }
        return l;
    }
};

// ---------------------------------------------------------------------------
// Kernel blocks.  Fsubmatrix (matroxTest-V1/compression.h:256-312): the
// squared distance is expanded as |a|^2 + |b|^2 - 2 a.b before exp(-r^2/(2h^2)),
// the same expansion the reference computes with dgemm + ddot.
static void kernel_block(const Handle &H, const int *amap, int lena, const int *bmap, int lenb,
                         double *out /* lena x lenb col-major */) {
    const int d = H.d;
    const double *X = H.X.data();
    std::vector<double> an(lena), bn(lenb);
    for (int i = 0; i < lena; ++i) {
        const double *a = X + (size_t)amap[i] * d;
        double s = 0;
        for (int p = 0; p < d; ++p) s += a[p] * a[p];
        an[i] = s;
    }
    for (int j = 0; j < lenb; ++j) {
        const double *b = X + (size_t)bmap[j] * d;
        double s = 0;
        for (int p = 0; p < d; ++p) s += b[p] * b[p];
        bn[j] = s;
    }
    const double kscal = -0.5 / (H.cfg.h * H.cfg.h);
    for (int j = 0; j < lenb; ++j) {
        const double *b = X + (size_t)bmap[j] * d;
        double *col = out + (size_t)j * lena;
        for (int i = 0; i < lena; ++i) {
            const double *a = X + (size_t)amap[i] * d;
            double dot = 0;
            for (int p = 0; p < d; ++p) dot += a[p] * b[p];
            col[i] = std::exp(kscal * (-2.0 * dot + an[i] + bn[j]));
        }
    }
}

// ---------------------------------------------------------------------------
// Cluster tree.  boundingbox::split (sympiler/boundingbox.cpp:337-447): project
// on the direction between the point farthest from the mean and the point
// farthest from that one, cut at the median, ties filling the left half up to
// ceil(num/2).  Children of node v are 2v+1 and 2v+2 (BFS numbering,
// ClusterTree.cpp:81-178).  The reference splits any node with more than m
// points; here the depth is the smallest one whose largest node has at most m
// points, which is the same tree whenever a level's node sizes agree, and keeps
// the tree perfect otherwise (the far-list search assumes a perfect tree,
// ClusterTree.cpp:665-724).
static void split_points(const Handle &H, const std::vector<int> &lids, std::vector<int> &left,
                         std::vector<int> &right) {
    const int d = H.d;
    const double *X = H.X.data();
    const int len = (int)lids.size();
    std::vector<double> center(d, 0.0);
    for (int i = 0; i < len; ++i)
        for (int p = 0; p < d; ++p) center[p] += X[(size_t)lids[i] * d + p];
    for (int p = 0; p < d; ++p) center[p] /= len;

    auto farthest_from = [&](const double *c) {
        double best = -1.0;
        int idx = 0;
        for (int i = 0; i < len; ++i) {
            double r = 0;
            for (int p = 0; p < d; ++p) {
                double t = X[(size_t)lids[i] * d + p] - c[p];
                r += t * t;
            }
            if (i == 0 || r > best) { best = r; idx = i; }
        }
        return idx;
    };
    int x0 = farthest_from(center.data());
    std::vector<double> p0(X + (size_t)lids[x0] * d, X + (size_t)lids[x0] * d + d);
    int x1 = farthest_from(p0.data());
    std::vector<double> dir(d);
    for (int p = 0; p < d; ++p) dir[p] = X[(size_t)lids[x1] * d + p] - p0[p];

    std::vector<double> proj(len, 0.0);
    for (int i = 0; i < len; ++i) {
        double s = 0;
        for (int p = 0; p < d; ++p) s += X[(size_t)lids[i] * d + p] * dir[p];
        proj[i] = s;
    }
    std::vector<double> sorted(proj);
    std::sort(sorted.begin(), sorted.end());
    double median = (sorted[len / 2] + sorted[len / 2 - 1]) / 2;
    size_t llength = (size_t)(len / 2 + (len % 2 != 0 ? 1 : 0));
    std::vector<int> middle;
    left.clear();
    right.clear();
    for (int i = 0; i < len; ++i) {
        if (proj[i] < median) left.push_back(lids[i]);
        else if (proj[i] > median) right.push_back(lids[i]);
        else middle.push_back(lids[i]);
    }
    for (int v : middle) {
        if (left.size() < llength) left.push_back(v);
        else right.push_back(v);
    }
}

static void box_of(const Handle &H, const std::vector<int> &lids, double *mid, double *radius) {
    const int d = H.d;
    std::vector<double> lo(d), hi(d);
    for (int p = 0; p < d; ++p) lo[p] = hi[p] = H.X[(size_t)lids[0] * d + p];
    for (size_t i = 1; i < lids.size(); ++i)
        for (int p = 0; p < d; ++p) {
            double x = H.X[(size_t)lids[i] * d + p];
            lo[p] = std::min(lo[p], x);
            hi[p] = std::max(hi[p], x);
        }
    double r = 0;
    for (int p = 0; p < d; ++p) {
        mid[p] = (lo[p] + hi[p]) / 2;
        double t = hi[p] - mid[p];
        r += t * t;
    }
    *radius = std::sqrt(r);
}

static void build_tree(Handle &H) {
    const int m = H.cfg.leaf_size;
    int depth = 0;
    while ((((int64_t)H.n + (1LL << depth) - 1) >> depth) > m) ++depth;
    if (depth < 1) throw std::runtime_error("tree needs at least two leaves: n must exceed leaf_size");
    if (depth > 20) throw std::runtime_error("tree deeper than 20 levels");
    H.depth = depth;
    H.nleaf = 1 << depth;
    H.numnodes = 2 * H.nleaf - 1;
    H.pts.assign(H.numnodes, {});
    H.pts[0].resize(H.n);
    for (int i = 0; i < H.n; ++i) H.pts[0][i] = i;
    for (int l = 0; l < depth; ++l) {
        int beg = (1 << l) - 1, end = (1 << (l + 1)) - 1;
#pragma omp parallel for schedule(dynamic)
        for (int v = beg; v < end; ++v) split_points(H, H.pts[v], H.pts[2 * v + 1], H.pts[2 * v + 2]);
    }
    for (int v = 0; v < H.numnodes; ++v)
        if (H.pts[v].empty()) throw std::runtime_error("empty tree node (too few points for the depth)");
    H.mid.assign((size_t)H.numnodes * H.d, 0.0);
    H.radius.assign(H.numnodes, 0.0);
#pragma omp parallel for schedule(dynamic)
    for (int v = 0; v < H.numnodes; ++v) box_of(H, H.pts[v], &H.mid[(size_t)v * H.d], &H.radius[v]);
    // sflevelsets: level 0 = leaves, ascending node ids (ClusterTree.cpp:196-232)
    H.sflevel.assign(depth + 1, {});
    for (int l = 0; l <= depth; ++l) {
        int dl = depth - l;
        for (int v = (1 << dl) - 1; v < (1 << (dl + 1)) - 1; ++v) H.sflevel[l].push_back(v);
    }
}

// Budget near lists (ClusterTree.cpp:508-554): each leaf keeps itself and the
// leaves with the smallest centre distance / (r_i + r_j), up to
// (int)(nleaf*budget - 1) entries; lists are then sorted (ClusterTree.cpp:63-74).
static void build_near(Handle &H) {
    const int L = H.nleaf, f = H.leaf_first(), d = H.d;
    H.nearn.assign(H.numnodes, {});
    const int cap = (int)((double)L * H.cfg.budget - 1);
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < L; ++i) {
        std::vector<std::pair<double, int>> ad(L);
        const double *mi = &H.mid[(size_t)(f + i) * d];
        for (int j = 0; j < L; ++j) {
            const double *mj = &H.mid[(size_t)(f + j) * d];
            double s = 0;
            for (int p = 0; p < d; ++p) s += (mi[p] - mj[p]) * (mi[p] - mj[p]);
            ad[j] = {std::sqrt(s) / (H.radius[f + i] + H.radius[f + j]), j};
        }
        ad[i].first = -1.0;  // the leaf itself always comes first
        std::sort(ad.begin(), ad.end());
        auto &nl = H.nearn[f + i];
        nl.push_back(f + i);
        for (int j = 1; j < L; ++j) {
            if ((int)nl.size() < cap) nl.push_back(f + ad[j].second);
            else break;
        }
        std::sort(nl.begin(), nl.end());
    }
}

// Far lists (ClusterTree.cpp:629-684): a leaf's far set is the set of maximal
// subtrees holding none of its near leaves (FindFarNodes from the root); then,
// bottom-up, far nodes common to both children move to the parent
// (MergeFarNodes(idx)).  Lists are sorted at the end (sortnodes).
static bool subtree_has_any(const Handle &H, int root, const std::vector<int> &nodes) {
    int rl = H.node_level(root);
    for (int v : nodes) {
        int vl = H.node_level(v);
        if (vl < rl) continue;
        int lo = root, hi = root;
        for (int i = rl + 1; i <= vl; ++i) { lo = 2 * lo + 1; hi = 2 * hi + 2; }
        if (v >= lo && v <= hi) return true;
    }
    return false;
}

static void find_far(const Handle &H, int root, int leaf, std::vector<int> &out) {
    if (subtree_has_any(H, root, H.nearn[leaf])) {
        if (!H.is_leaf(root)) {
            find_far(H, 2 * root + 1, leaf, out);
            find_far(H, 2 * root + 2, leaf, out);
        }
    } else {
        out.push_back(root);
    }
}

static void build_far(Handle &H) {
    H.farn.assign(H.numnodes, {});
    const int f = H.leaf_first();
#pragma omp parallel for schedule(dynamic)
    for (int v = f; v < H.numnodes; ++v) find_far(H, 0, v, H.farn[v]);
    for (int l = 1; l < H.depth; ++l) {
        for (int v : H.sflevel[l]) {
            auto &a = H.farn[2 * v + 1];
            auto &b = H.farn[2 * v + 2];
            std::set<int> sb(b.begin(), b.end());
            std::vector<int> common, na;
            for (int w : a) (sb.count(w) ? common : na).push_back(w);
            if (common.empty()) continue;
            std::set<int> sc(common.begin(), common.end());
            std::vector<int> nb;
            for (int w : b)
                if (!sc.count(w)) nb.push_back(w);
            a.swap(na);
            b.swap(nb);
            for (int w : common) H.farn[v].push_back(w);
        }
    }
    for (int v = 1; v < H.numnodes; ++v) std::sort(H.farn[v].begin(), H.farn[v].end());
    H.is_hss = true;
    for (int v = f; v < H.numnodes; ++v)
        if (H.nearn[v].size() != 1) H.is_hss = false;
}

// Coarsened level sets: parlevel(agg, llevel) then genpostw
// (ClusterTree.cpp:732-768, 845-857).  Group 0 holds the bottom `llevel`
// levels, later groups `agg` levels each; a partition is the post-order of a
// subtree cut to the group's levels.  The root level is never included.
static void postorder(const Handle &H, std::vector<int> &out, int v, int clevel, int elevel) {
    if (clevel == elevel) { out.push_back(v); return; }
    postorder(H, out, 2 * v + 1, clevel - 1, elevel);
    postorder(H, out, 2 * v + 2, clevel - 1, elevel);
    out.push_back(v);
}

static void build_coarsening(Handle &H) {
    const int agg = std::max(1, H.cfg.agg_levels);
    const int nlev = H.depth;  // sflevel indices 0..depth-1 (root excluded)
    const int llevel = std::min(std::max(1, H.cfg.bottom_levels), nlev);
    H.coarlevels.clear();
    H.coarlevels.push_back({});
    for (int i = 0; i < llevel; ++i) H.coarlevels[0].push_back(i);
    for (int i = llevel; i < nlev; ++i) {
        if (H.coarlevels.back().size() == (size_t)agg || H.coarlevels.size() == 1)
            H.coarlevels.push_back({});
        H.coarlevels.back().push_back(i);
    }
    H.postw.assign(H.coarlevels.size(), {});
    for (size_t g = 0; g < H.coarlevels.size(); ++g) {
        int tl = H.coarlevels[g].back(), el = H.coarlevels[g].front();
        for (int v : H.sflevel[tl]) {
            H.postw[g].push_back({});
            postorder(H, H.postw[g].back(), v, tl, el);
        }
    }
}

// BalanceCoarLevelSet + binpacking (matroxTest-V1/compression.h:953-1053):
// partitions of a group are sorted by decreasing transfer cost and dealt to
// the currently cheapest bin.  Only the memory order of VT and the skeleton
// buffers depends on it (hcds.h:289-332, 362-393).  Equal costs keep their
// partition order (stable sort) so the layout is reproducible.
static void balance_partitions(Handle &H) {
    H.opostw.assign(H.postw.size(), {});
    for (size_t g = 0; g < H.postw.size(); ++g) {
        auto &parts = H.postw[g];
        int nw = (int)parts.size();
        int nbins = H.cfg.nbins > 0 ? H.cfg.nbins : 1;
        int nparts = nw >= nbins ? nbins : nw / 2;
        if (nparts == 0) nparts = 1;
        std::vector<std::pair<uint64_t, int>> cost(nw);
        for (int i = 0; i < nw; ++i) {
            uint64_t c = 0;
            for (int v : parts[i]) {
                if (H.is_leaf(v)) c += 2ULL * H.skels[v].size() * H.projcol[v];
                else
                    c += 2ULL * H.skels[v].size() * (H.skels[2 * v + 1].size() + H.skels[2 * v + 2].size());
            }
            cost[i] = {c, i};
        }
        std::stable_sort(cost.begin(), cost.end(),
                         [](const std::pair<uint64_t, int> &a, const std::pair<uint64_t, int> &b) {
                             return a.first > b.first;
                         });
        std::vector<uint64_t> load(nparts, 0);
        H.opostw[g].assign(nparts, {});
        for (int i = 0; i < nw; ++i) {
            int mb = 0;
            uint64_t mv = std::numeric_limits<uint64_t>::max();
            for (int b = 0; b < nparts; ++b)
                if (load[b] < mv) { mv = load[b]; mb = b; }
            load[mb] += cost[i].first;
            auto &src = parts[cost[i].second];
            H.opostw[g][mb].insert(H.opostw[g][mb].end(), src.begin(), src.end());
        }
    }
}

// ---------------------------------------------------------------------------
// Approximate k nearest neighbours by randomized projection trees (the scheme
// of sympiler/NearSearch.cpp:20-60 and NNtree.cpp:13-185): up to 10 trees,
// leaf size 4k doubling (below 2048) while the recall measured on 10 test
// points stays under 0.8.  Every point's list is kept sorted by distance, so
// the nearest half (the pruning neighbours of compression.h:362-377) and the
// farther half (the sampling candidates) are well defined.
static void build_knn(Handle &H) {
    const int n = H.n, d = H.d, k = std::min(H.cfg.knn_k, n);
    H.knn_k = k;
    const double *X = H.X.data();
    H.NN.assign((size_t)n * k, {std::numeric_limits<double>::max(), -1});
    auto dist2 = [&](int a, int b) {
        double s = 0;
        for (int p = 0; p < d; ++p) {
            double t = X[(size_t)a * d + p] - X[(size_t)b * d + p];
            s += t * t;
        }
        return s;
    };
    // per-point max-heaps with membership sets
    auto heap_insert = [&](int i, double dd, int j) {
        auto *h = &H.NN[(size_t)i * k];
        if (dd >= h[0].first) return;
        for (int t = 0; t < k; ++t)
            if (h[t].second == j) return;
        std::pop_heap(h, h + k);
        h[k - 1] = {dd, j};
        std::push_heap(h, h + k);
    };
    for (int i = 0; i < n; ++i) std::make_heap(&H.NN[(size_t)i * k], &H.NN[(size_t)i * k] + k);
    int mleaf = std::max(4 * k, 32);
    const int ntest = std::min(10, n);
    // exact neighbours of the test points
    std::vector<std::vector<int>> exact(ntest);
    for (int t = 0; t < ntest; ++t) {
        std::vector<std::pair<double, int>> all(n);
        for (int j = 0; j < n; ++j) all[j] = {dist2(t, j), j};
        std::partial_sort(all.begin(), all.begin() + k, all.end());
        for (int j = 0; j < k; ++j) exact[t].push_back(all[j].second);
    }
    for (int tree = 0; tree < 10; ++tree) {
        int depth = (int)(std::log((double)n / mleaf) / std::log(2.0));
        if (depth < 0) depth = 0;
        std::vector<std::vector<int>> lev(1);
        lev[0].resize(n);
        for (int i = 0; i < n; ++i) lev[0][i] = i;
        for (int l = 0; l < depth; ++l) {
            std::vector<std::vector<int>> nxt(lev.size() * 2);
#pragma omp parallel for schedule(dynamic)
            for (int b = 0; b < (int)lev.size(); ++b) {
                auto &ids = lev[b];
                std::vector<double> dir(d);
                for (int p = 0; p < d; p += 2) {
                    double z0, z1;
                    normal_pair(stream_seed(stream_seed(H.cfg.seed, 0x4E4E, tree), l, b), p / 2, &z0, &z1);
                    dir[p] = z0;
                    if (p + 1 < d) dir[p + 1] = z1;
                }
                std::vector<double> pr(ids.size());
                for (size_t i = 0; i < ids.size(); ++i) {
                    double s = 0;
                    for (int p = 0; p < d; ++p) s += X[(size_t)ids[i] * d + p] * dir[p];
                    pr[i] = s;
                }
                std::vector<double> srt(pr);
                std::nth_element(srt.begin(), srt.begin() + srt.size() / 2, srt.end());
                double med = srt[srt.size() / 2];
                auto &L = nxt[2 * b], &R = nxt[2 * b + 1];
                std::vector<int> mid;
                for (size_t i = 0; i < ids.size(); ++i) {
                    if (pr[i] < med) L.push_back(ids[i]);
                    else if (pr[i] > med) R.push_back(ids[i]);
                    else mid.push_back(ids[i]);
                }
                for (int v : mid) (L.size() <= R.size() ? L : R).push_back(v);
            }
            lev.swap(nxt);
        }
#pragma omp parallel for schedule(dynamic)
        for (int b = 0; b < (int)lev.size(); ++b) {
            auto &ids = lev[b];
            for (int i : ids)
                for (int j : ids) heap_insert(i, dist2(i, j), j);
        }
        double acc = 0;
        for (int t = 0; t < ntest; ++t) {
            std::set<int> ex(exact[t].begin(), exact[t].end());
            int c = 0;
            for (int j = 0; j < k; ++j)
                if (ex.count(H.NN[(size_t)t * k + j].second)) ++c;
            acc += (double)c / k;
        }
        acc /= ntest;
        if (acc >= 0.8) break;
        if (2 * mleaf < 2048) mleaf *= 2;
    }
#pragma omp parallel for
    for (int i = 0; i < n; ++i) std::sort(&H.NN[(size_t)i * k], &H.NN[(size_t)i * k] + k);
}

// ---------------------------------------------------------------------------
// Interpolative decomposition (decomposition, compression.h:446-523):
// Householder QR with column pivoting of the sampled block A (m x n,
// col-major).  The rank s is the first j >= 1 with |R_jj / R_00| < tol, or n,
// capped at max_rank; proj (s x n) = R11^{-1} [R11 R12] with columns put back
// in original order, so K(:, cols) ~= K(:, skels) * proj.  The factorization
// stops once s is known: the trailing rows are never needed.
static int interpolative_decomposition(double *A, int m, int n, double tol, int max_rank,
                                       std::vector<int> &skel_cols, std::vector<double> &proj,
                                       bool inner_parallel, double *left_out_ratio) {
    *left_out_ratio = 0.0;
    std::vector<int> jpvt(n);
    for (int j = 0; j < n; ++j) jpvt[j] = j;
    std::vector<double> vn1(n), vn2(n);
    for (int j = 0; j < n; ++j) {
        double s = 0;
        for (int i = 0; i < m; ++i) s += A[(size_t)j * m + i] * A[(size_t)j * m + i];
        vn1[j] = vn2[j] = std::sqrt(s);
    }
    const double tol3z = std::sqrt(std::numeric_limits<double>::epsilon());
    const int kmax = std::min(m, n);
    int s = std::min(n, max_rank);
    double r00 = 0.0;
    bool r00_zero = false;
    std::vector<double> v(m);
    for (int j = 0; j < kmax; ++j) {
        // pivot
        int p = j;
        for (int t = j + 1; t < n; ++t)
            if (vn1[t] > vn1[p]) p = t;
        if (p != j) {
            for (int i = 0; i < m; ++i) std::swap(A[(size_t)p * m + i], A[(size_t)j * m + i]);
            std::swap(jpvt[p], jpvt[j]);
            vn1[p] = vn1[j];
            vn2[p] = vn2[j];
        }
        // Householder vector for A[j:m, j]
        double *aj = A + (size_t)j * m;
        double alpha = aj[j], xnorm = 0;
        for (int i = j + 1; i < m; ++i) xnorm += aj[i] * aj[i];
        xnorm = std::sqrt(xnorm);
        double beta_h = 0.0, rjj;
        if (xnorm == 0.0) {
            rjj = alpha;
        } else {
            double nrm = std::sqrt(alpha * alpha + xnorm * xnorm);
            rjj = alpha >= 0 ? -nrm : nrm;
            beta_h = (rjj - alpha) / rjj;  // tau
            double scal = 1.0 / (alpha - rjj);
            for (int i = j + 1; i < m; ++i) aj[i] *= scal;
        }
        aj[j] = rjj;
        if (j == 0) {
            r00 = rjj;
            r00_zero = (r00 == 0.0);
        } else if (!r00_zero && std::abs(rjj / r00) < tol) {
            s = std::min(j, max_rank);
            *left_out_ratio = std::abs(rjj / r00);
            break;
        }
        if (j + 1 >= std::min(n, max_rank)) {
            // rows 0..j of R are complete for all columns once the trailing
            // columns are updated by this reflection
            s = std::min(n, max_rank);
        }
        // apply H_j = I - tau v v^T (v = [1; aj[j+1:m]]) to trailing columns
        if (beta_h != 0.0) {
#pragma omp parallel for if (inner_parallel) schedule(static)
            for (int c = j + 1; c < n; ++c) {
                double *ac = A + (size_t)c * m;
                double w = ac[j];
                for (int i = j + 1; i < m; ++i) w += aj[i] * ac[i];
                w *= beta_h;
                ac[j] -= w;
                for (int i = j + 1; i < m; ++i) ac[i] -= w * aj[i];
            }
        }
        if (j + 1 >= std::min(n, max_rank)) {
            // the cap stops here: the next pivot would be the largest
            // trailing column norm of rows j+1..m
            if (j + 1 < n && j + 1 < m && !r00_zero) {
                double mx = 0.0;
                for (int c = j + 1; c < n; ++c) {
                    double ss = 0;
                    for (int i = j + 1; i < m; ++i) ss += A[(size_t)c * m + i] * A[(size_t)c * m + i];
                    mx = std::max(mx, ss);
                }
                *left_out_ratio = std::sqrt(mx) / std::abs(r00);
            }
            break;
        }
        // downdate column norms (LAPACK dlaqp2 recipe)
        for (int c = j + 1; c < n; ++c) {
            if (vn1[c] != 0.0) {
                double temp = std::abs(A[(size_t)c * m + j]) / vn1[c];
                temp = 1.0 - temp * temp;
                temp = temp < 0 ? 0 : temp;
                double temp2 = temp * (vn1[c] / vn2[c]) * (vn1[c] / vn2[c]);
                if (temp2 <= tol3z) {
                    double ss = 0;
                    for (int i = j + 1; i < m; ++i) ss += A[(size_t)c * m + i] * A[(size_t)c * m + i];
                    vn1[c] = std::sqrt(ss);
                    vn2[c] = vn1[c];
                } else {
                    vn1[c] *= std::sqrt(temp);
                }
            }
        }
    }
    if (s > n) s = n;
    if (s < 1) s = 1;
    skel_cols.assign(jpvt.begin(), jpvt.begin() + s);
    proj.assign((size_t)s * n, 0.0);
    if (r00_zero) return s;  // zero block: proj stays zero (compression.h:499)
    // T = R11^{-1} R(0:s, :) by back substitution, columns in pivoted order
    std::vector<double> T((size_t)s * n);
    for (int c = 0; c < n; ++c) {
        double *t = &T[(size_t)c * s];
        for (int i = 0; i < s; ++i) t[i] = (i <= c) ? A[(size_t)c * m + i] : 0.0;
        for (int i = s - 1; i >= 0; --i) {
            double x = t[i];
            for (int q = i + 1; q < s; ++q) x -= A[(size_t)q * m + i] * t[q];
            t[i] = x / A[(size_t)i * m + i];
        }
    }
    for (int c = 0; c < n; ++c)
        for (int i = 0; i < s; ++i) proj[(size_t)jpvt[c] * s + i] = T[(size_t)c * s + i];
    return s;
}

// Sampling rows and skeletonising one node (skeletonize_leaf and
// skeletonize_internal, compression.h:597-797 with BuildNeighBorsLeaf/Internal
// :362-443).  Rows: the nearest sampled neighbours first, then uniform random
// points outside the node (a per-node random stream replaces rand()); when the
// node leaves too few outside points, all n points are used.
struct NodeSampling {
    std::unordered_set<int> pnids;   // pruning neighbours
    std::map<int, double> snids;     // sampling neighbours -> distance
};

static void skeletonize(Handle &H, int v, std::vector<NodeSampling> &samp, bool inner_parallel) {
    const int n = H.n, k = H.knn_k, m = H.cfg.leaf_size;
    const bool leaf = H.is_leaf(v);
    std::vector<int> bmap;
    if (leaf) bmap = H.pts[v];
    else {
        bmap = H.skels[2 * v + 1];
        bmap.insert(bmap.end(), H.skels[2 * v + 2].begin(), H.skels[2 * v + 2].end());
    }
    const auto &own = H.pts[v];
    std::unordered_set<int> ownset(own.begin(), own.end());
    size_t nsamples = 2 * bmap.size();
    if (nsamples < 2 * (size_t)m) nsamples = 2 * (size_t)m;
    auto &S = samp[v];
    if (leaf) {
        S.pnids.clear();
        for (int ii = 0; ii < k / 2; ++ii)
            for (int p : own) S.pnids.insert(H.NN[(size_t)p * k + ii].second);
        for (int p : own) S.pnids.erase(p);
        std::vector<std::pair<double, int>> tmp;
        for (int ii = (k + 1) / 2; ii < k; ++ii)
            for (int p : own) tmp.push_back(H.NN[(size_t)p * k + ii]);
        std::sort(tmp.begin(), tmp.end());
        S.snids.clear();
        size_t lim = (size_t)(k - 1) * own.size() / 2;
        for (size_t i = 0; i < tmp.size() && i < lim && S.snids.size() < nsamples; ++i) {
            int q = tmp[i].second;
            if (q < 0) continue;
            if (!S.pnids.count(q) && !ownset.count(q)) S.snids.insert({q, tmp[i].first});
        }
    } else {
        auto &Ls = samp[2 * v + 1], &Rs = samp[2 * v + 2];
        S.snids = Ls.snids;
        for (auto &e : Rs.snids) {
            auto r = S.snids.insert(e);
            if (!r.second && r.first->second > e.second) r.first->second = e.second;
        }
        for (int p : own) S.snids.erase(p);
        for (int p : Ls.pnids) S.snids.erase(p);
        for (int p : Rs.pnids) S.snids.erase(p);
    }
    std::vector<std::pair<double, int>> ordered;
    for (auto &e : S.snids) ordered.push_back({e.second, e.first});
    std::sort(ordered.begin(), ordered.end());

    std::vector<int> amap;
    if (!H.cfg.all_rows && nsamples < (size_t)(n - (int)own.size())) {
        std::unordered_set<int> inA;
        for (size_t i = 0; i < ordered.size() && amap.size() < nsamples; ++i) {
            amap.push_back(ordered[i].second);
            inA.insert(ordered[i].second);
        }
        SplitMix64 g(stream_seed(H.cfg.seed, 0x5EED, (uint64_t)v));
        while (amap.size() < nsamples) {
            int q = (int)g.below((uint64_t)n);
            if (!inA.count(q) && !ownset.count(q)) {
                amap.push_back(q);
                inA.insert(q);
            }
        }
    } else {
        amap.resize(n);
        for (int q = 0; q < n; ++q) amap[q] = q;
    }
    const int m1 = (int)amap.size(), n1 = (int)bmap.size();
    std::vector<double> Kab((size_t)m1 * n1);
    kernel_block(H, amap.data(), m1, bmap.data(), n1, Kab.data());
    std::vector<int> cols;
    std::vector<double> proj;
    interpolative_decomposition(Kab.data(), m1, n1, H.cfg.tol, H.cfg.max_rank, cols, proj, inner_parallel,
                                &H.id_ratio[v]);
    H.id_rows[v] = m1;
    H.skels[v].resize(cols.size());
    for (size_t i = 0; i < cols.size(); ++i) H.skels[v][i] = bmap[cols[i]];
    H.proj[v].swap(proj);
    H.projcol[v] = n1;
    S.pnids.clear();
    for (int p : H.skels[v])
        for (int jj = 0; jj < k / 2; ++jj) S.pnids.insert(H.NN[(size_t)p * k + jj].second);
}

// compression() (compression.h:1057-1076): leaves in parallel, then each
// level bottom-up.
static void compress(Handle &H) {
    H.skels.assign(H.numnodes, {});
    H.proj.assign(H.numnodes, {});
    H.projcol.assign(H.numnodes, 0);
    H.id_ratio.assign(H.numnodes, 0.0);
    H.id_rows.assign(H.numnodes, 0);
    std::vector<NodeSampling> samp(H.numnodes);
    const int nthr = omp_get_max_threads();
    for (int l = 0; l < H.depth; ++l) {
        auto &lev = H.sflevel[l];
        bool inner = (int)lev.size() < nthr;
        if (inner) {
            for (int v : lev) skeletonize(H, v, samp, true);
        } else {
#pragma omp parallel for schedule(dynamic)
            for (int i = 0; i < (int)lev.size(); ++i) skeletonize(H, lev[i], samp, false);
        }
    }
}

// ---------------------------------------------------------------------------
// Flattened structure arrays (sympiler/HTree.cpp:299-457) and blocking
// (ClusterTree.cpp:951-1089, HTree.cpp:88-165).
static void emit_structure(Handle &H) {
    const int nn = H.numnodes, L = H.nleaf, f = H.leaf_first();
    auto &I = H.i32;
    // levelset / idx: top-down levels without the root (HTree.cpp:351-365)
    {
        auto &ls = I["levelset"];
        auto &ix = I["idx_level"];
        ls.assign(1, 0);
        for (int l = H.depth - 1; l >= 0; --l) {
            for (int v : H.sflevel[l]) ix.push_back(v);
            ls.push_back((int)ix.size());
        }
    }
    // coarsened: clevelset / wpart / idx (cidx), HTree.cpp:302-333
    {
        auto &cl = I["clevelset"];
        auto &wp = I["wpart"];
        auto &cx = I["idx"];
        cl.push_back(0);
        wp.push_back(0);
        for (auto &grp : H.postw) {
            for (auto &part : grp) {
                for (int v : part) cx.push_back(v);
                wp.push_back((int)cx.size());
            }
            cl.push_back((int)wp.size() - 1);
        }
    }
    auto &lc = I["lchildren"];
    auto &rc = I["rchildren"];
    lc.assign(nn, -1);
    rc.assign(nn, -1);
    for (int v = 0; v < f; ++v) { lc[v] = 2 * v + 1; rc[v] = 2 * v + 2; }
    auto &lm = I["lm"];
    lm.assign(nn, -1);
    auto &leaf = I["leaf"];
    leaf.resize(L);
    auto &Ddim = I["Ddim"];
    Ddim.resize(L);
    for (int i = 0; i < L; ++i) {
        lm[f + i] = i;
        leaf[i] = f + i;
        Ddim[i] = (int)H.pts[f + i].size();
    }
    // lids: leaves first, then each level upward, root excluded (HTree.cpp:429-456)
    auto &lids = I["lids"];
    auto &lo = I["lidsoffset"];
    auto &ll = I["lidslen"];
    lo.assign(nn, 0);
    ll.assign(nn, 0);
    for (int l = 0; l < H.depth; ++l)
        for (int v : H.sflevel[l]) {
            lo[v] = (int)lids.size();
            ll[v] = (int)H.pts[v].size();
            lids.insert(lids.end(), H.pts[v].begin(), H.pts[v].end());
        }
    // flat near / far pairs (StructureNear/StructureFar, HTree.cpp:168-223)
    auto &fnx = I["flat_nx"], &fny = I["flat_ny"];
    for (int i = 0; i < L; ++i)
        for (int w : H.nearn[f + i]) { fnx.push_back(i); fny.push_back(w - f); }
    auto &ffx = I["flat_fx"], &ffy = I["flat_fy"];
    for (int v = 1; v < nn; ++v)
        for (int w : H.farn[v]) { ffx.push_back(v); ffy.push_back(w); }
    // near blocking NearBlock(bs): row-block = bs leaves; column blocks sorted
    {
        const int bs = std::max(1, H.cfg.near_block);
        const int nb = (L - 1 + bs) / bs;
        std::vector<std::vector<std::vector<std::pair<int, int>>>> blocks(nb);
        for (int r = 0; r < nb; ++r) {
            std::vector<int> cols;
            std::vector<char> used(nb, 0);
            for (int x = r * bs; x < std::min((r + 1) * bs, L); ++x)
                for (int w : H.nearn[f + x]) {
                    int cb = (w - f) / bs;
                    if (!used[cb]) { used[cb] = 1; cols.push_back(cb); }
                }
            std::sort(cols.begin(), cols.end());
            std::map<int, int> pos;
            for (size_t j = 0; j < cols.size(); ++j) pos[cols[j]] = (int)j;
            blocks[r].resize(cols.size());
            for (int x = r * bs; x < std::min((r + 1) * bs, L); ++x)
                for (int w : H.nearn[f + x]) blocks[r][pos[(w - f) / bs]].push_back({x, w - f});
        }
        auto &nbs = I["nblockSet"], &nbk = I["nblocks"], &px = I["npairx"], &py = I["npairy"];
        nbs.push_back(0);
        nbk.push_back(0);
        for (auto &row : blocks) {
            for (auto &cb : row) {
                for (auto &pr : cb) { px.push_back(pr.first); py.push_back(pr.second); }
                nbk.push_back((int)px.size());
            }
            nbs.push_back((int)nbk.size() - 1);
        }
        H.scal["nl"] = nb;
    }
    // far blocking FarBlock(bs, false): nodes in bottom-up level order (nodeId)
    {
        const int bs = std::max(1, H.cfg.far_block);
        std::vector<int> nodeId;
        std::vector<int> mapIdx(nn, -1);
        for (int l = 0; l < H.depth; ++l)
            for (int v : H.sflevel[l]) { mapIdx[v] = (int)nodeId.size(); nodeId.push_back(v); }
        const int Nn = nn - 1;
        const int fb = (Nn - 1 + bs) / bs;
        std::vector<std::vector<std::vector<std::pair<int, int>>>> blocks(fb);
        for (int r = 0; r < fb; ++r) {
            std::vector<int> cols;
            std::vector<char> used(fb, 0);
            for (int i = r * bs; i < std::min((r + 1) * bs, (int)nodeId.size()); ++i)
                for (int w : H.farn[nodeId[i]]) {
                    int cb = mapIdx[w] / bs;
                    if (!used[cb]) { used[cb] = 1; cols.push_back(cb); }
                }
            std::sort(cols.begin(), cols.end());
            std::map<int, int> pos;
            for (size_t j = 0; j < cols.size(); ++j) pos[cols[j]] = (int)j;
            blocks[r].resize(cols.size());
            for (int i = r * bs; i < std::min((r + 1) * bs, (int)nodeId.size()); ++i)
                for (int w : H.farn[nodeId[i]]) blocks[r][pos[mapIdx[w] / bs]].push_back({nodeId[i], w});
        }
        auto &fbs = I["fblockSet"], &fbk = I["fblocks"], &px = I["fpairx"], &py = I["fpairy"];
        fbs.push_back(0);
        fbk.push_back(0);
        for (auto &row : blocks) {
            for (auto &cb : row) {
                for (auto &pr : cb) { px.push_back(pr.first); py.push_back(pr.second); }
                fbk.push_back((int)px.size());
            }
            fbs.push_back((int)fbk.size() - 1);
        }
        H.scal["fl"] = fb;
    }
    H.scal["n"] = H.n;
    H.scal["d"] = H.d;
    H.scal["nleaf"] = L;
    H.scal["numnodes"] = nn;
    H.scal["depth"] = H.depth;
    H.scal["cdepth"] = (int64_t)H.postw.size();
    H.scal["ncount"] = (int64_t)I["flat_nx"].size();
    H.scal["fcount"] = (int64_t)I["flat_fx"].size();
    H.scal["is_hss"] = H.is_hss ? 1 : 0;
    H.scal["N_total"] = H.n;
}

// CDS packing (DUV2CDS, matroxTest-V1/hcds.h:433-460): near blocks D_k =
// K(lids_x, lids_y) and far blocks B_k = K(skels_x, skels_y) in pair order;
// VT blocks (proj, s x cols) and skeleton-buffer offsets in bin-packed
// post-order (coarcacheProj / coarcacheWUskel); per-leaf W/Y offsets
// (cachewuoffset); skeleton lengths with slen[0] = 0 (cacheSkeldim).
static void emit_cds(Handle &H) {
    auto &I = H.i32;
    const int nn = H.numnodes, L = H.nleaf, f = H.leaf_first();
    auto pack = [&](const std::vector<int32_t> &px, const std::vector<int32_t> &py, bool near,
                    std::vector<uint64_t> &ptr, std::vector<double> &buf) {
        const size_t cnt = px.size();
        ptr.resize(cnt);
        uint64_t off = 0;
        for (size_t k = 0; k < cnt; ++k) {
            ptr[k] = off;
            uint64_t a = near ? H.pts[f + px[k]].size() : H.skels[px[k]].size();
            uint64_t b = near ? H.pts[f + py[k]].size() : H.skels[py[k]].size();
            off += a * b;
        }
        H.scal[near ? "D_len" : "B_len"] = (int64_t)off;
        if (H.cfg.skip_blocks) {
            buf.clear();
            return;
        }
        buf.assign(off, 0.0);
#pragma omp parallel for schedule(dynamic)
        for (long long k = 0; k < (long long)cnt; ++k) {
            const std::vector<int> &a = near ? H.pts[f + px[k]] : H.skels[px[k]];
            const std::vector<int> &b = near ? H.pts[f + py[k]] : H.skels[py[k]];
            if (!a.empty() && !b.empty())
                kernel_block(H, a.data(), (int)a.size(), b.data(), (int)b.size(), buf.data() + ptr[k]);
        }
    };
    if (H.is_hss) {
        pack(I["flat_nx"], I["flat_ny"], true, H.u64["Dptr"], H.f64["D"]);
        pack(I["flat_fx"], I["flat_fy"], false, H.u64["Bptr"], H.f64["B"]);
        // the HSS evaluator indexes near blocks by leaf and far blocks by node-1
        // (flat order, HTree.cpp:168-223); the blocked arrays then describe the
        // same order as one row-block each, so the H2-b signature stays valid
        I["npairx"] = I["flat_nx"];
        I["npairy"] = I["flat_ny"];
        I["fpairx"] = I["flat_fx"];
        I["fpairy"] = I["flat_fy"];
        I["nblockSet"] = {0, 1};
        I["nblocks"] = {0, (int32_t)I["flat_nx"].size()};
        I["fblockSet"] = {0, 1};
        I["fblocks"] = {0, (int32_t)I["flat_fx"].size()};
        H.scal["nl"] = 1;
        H.scal["fl"] = 1;
    } else {
        pack(I["npairx"], I["npairy"], true, H.u64["Dptr"], H.f64["D"]);
        pack(I["fpairx"], I["fpairy"], false, H.u64["Bptr"], H.f64["B"]);
    }
    auto &vtp = I["VTptr"];
    auto &skr = I["skelrow"];
    vtp.assign(nn, 0);
    skr.assign(nn, 0);
    int64_t voff = 0, soff = 0;
    std::vector<int> order;
    for (auto &grp : H.opostw)
        for (auto &bin : grp)
            for (int v : bin) order.push_back(v);
    for (int v : order) {
        vtp[v] = (int32_t)voff;
        skr[v] = (int32_t)soff;
        voff += (int64_t)H.skels[v].size() * H.projcol[v];
        soff += (int64_t)H.skels[v].size();
    }
    if (voff > INT32_MAX) throw std::runtime_error("VT exceeds int32 offsets of the reference layout");
    auto &VT = H.f64["VT"];
    VT.assign(voff, 0.0);
    for (int v : order)
        std::copy(H.proj[v].begin(), H.proj[v].end(), VT.begin() + vtp[v]);
    auto &slen = I["slen"];
    auto &pc = I["proj_column"];
    slen.assign(nn, 0);
    pc.assign(nn, 0);
    for (int v = 1; v < nn; ++v) {
        slen[v] = (int)H.skels[v].size();
        pc[v] = H.projcol[v];
    }
    auto &wr = I["wrow"];
    wr.resize(L);
    int64_t r = 0;
    for (int i = 0; i < L; ++i) { wr[i] = (int32_t)r; r += H.pts[f + i].size(); }
    I["urow"] = wr;
    auto &sk = I["skels"];
    auto &sko = I["skelsoffset"];
    sko.assign(nn, 0);
    for (int v = 0; v < nn; ++v) {
        sko[v] = (int)sk.size();
        sk.insert(sk.end(), H.skels[v].begin(), H.skels[v].end());
    }
    H.f64["id_ratio"] = H.id_ratio;
    I["id_rows"].assign(H.id_rows.begin(), H.id_rows.end());
    H.scal["skel_rows"] = soff;
    H.scal["ncount"] = (int64_t)H.u64["Dptr"].size();
    H.scal["fcount"] = (int64_t)H.u64["Bptr"].size();
}

static void validate(const mxh_config &c, int n, int d) {
    if (n < 2 || d < 1) throw std::runtime_error("need n >= 2 points in d >= 1 dimensions");
    if (c.leaf_size < 1) throw std::runtime_error("leaf_size must be >= 1");
    if (c.max_rank < 1) throw std::runtime_error("max_rank must be >= 1");
    if (!(c.tol >= 0)) throw std::runtime_error("tol must be >= 0");
    if (!(c.budget >= 0 && c.budget <= 1)) throw std::runtime_error("budget must be in [0, 1]");
    if (!(c.h > 0)) throw std::runtime_error("bandwidth h must be > 0");
    if (c.kernel != MXH_KERNEL_GAUSSIAN) throw std::runtime_error("only the Gaussian kernel is supported");
    if (c.knn_k < 2) throw std::runtime_error("knn_k must be >= 2");
}

static Handle *build(const double *X, int n, int d, const mxh_config *cfg, bool full) {
    std::unique_ptr<Handle> H(new Handle());
    H->cfg = *cfg;
    validate(H->cfg, n, d);
    int saved = omp_get_max_threads();
    if (cfg->nthreads > 0) omp_set_num_threads(cfg->nthreads);
    try {
        H->n = n;
        H->d = d;
        H->X.assign(X, X + (size_t)n * d);
        H->f64["X"] = H->X;
        build_tree(*H);
        build_near(*H);
        build_far(*H);
        build_coarsening(*H);
        build_knn(*H);
        emit_structure(*H);
        if (full) {
            compress(*H);
            balance_partitions(*H);
            emit_cds(*H);
        }
    } catch (...) {
        omp_set_num_threads(saved);
        throw;
    }
    omp_set_num_threads(saved);
    return H.release();
}


// ---------------------------------------------------------------------------
// The reference's on-disk structure: raw int32 (and double) arrays, one file
// per array, written by savetree2disk / SaveSampling (sympiler/HTree.cpp:
// 234-249, 469-542) and read by readHtree (matroxTest-V1/HTree.h:76-260).
static void write_raw(const std::string &path, const void *data, size_t bytes) {
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + path);
    size_t w = bytes ? fwrite(data, 1, bytes, f) : 0;
    fclose(f);
    if (w != bytes) throw std::runtime_error("short write to " + path);
}

template <class T>
static bool read_raw(const std::string &path, std::vector<T> &out) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return false;
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    if (sz < 0 || sz % (long)sizeof(T) != 0) {
        fclose(f);
        throw std::runtime_error(path + ": size is not a whole number of elements");
    }
    out.resize((size_t)sz / sizeof(T));
    size_t r = sz ? fread(out.data(), 1, (size_t)sz, f) : 0;
    fclose(f);
    if (r != (size_t)sz) throw std::runtime_error("short read from " + path);
    return true;
}

static void save_bin(const Handle &H, const std::string &dir) {
    auto W = [&](const char *name, const std::vector<int32_t> &v) {
        write_raw(dir + "/" + name, v.data(), v.size() * sizeof(int32_t));
    };
    auto &I = H.i32;
    W("levelset.bin", I.at("levelset"));
    W("idx.bin", I.at("idx_level"));
    W("clevelset.bin", I.at("clevelset"));
    W("wpart.bin", I.at("wpart"));
    W("cidx.bin", I.at("idx"));
    W("lchildren.bin", I.at("lchildren"));
    W("rchildren.bin", I.at("rchildren"));
    if (!H.is_hss) {
        W("nblockset.bin", I.at("nblockSet"));
        W("nblocks.bin", I.at("nblocks"));
        W("nxval.bin", I.at("npairx"));
        W("nyval.bin", I.at("npairy"));
        W("fblockset.bin", I.at("fblockSet"));
        W("fblocks.bin", I.at("fblocks"));
        W("fxval.bin", I.at("fpairx"));
        W("fyval.bin", I.at("fpairy"));
    } else {
        W("nxval.bin", I.at("flat_nx"));
        W("nyval.bin", I.at("flat_ny"));
        W("fxval.bin", I.at("flat_fx"));
        W("fyval.bin", I.at("flat_fy"));
    }
    W("dim.bin", I.at("Ddim"));
    W("leafmap.bin", I.at("lm"));
    W("lids.bin", I.at("lids"));
    W("lidsoffset.bin", I.at("lidsoffset"));
    W("lidslen.bin", I.at("lidslen"));
    std::vector<int32_t> nn(H.NN.size());
    std::vector<double> dist(H.NN.size());
    for (size_t i = 0; i < H.NN.size(); ++i) {
        nn[i] = H.NN[i].second;
        dist[i] = H.NN[i].first;
    }
    W("NN.bin", nn);
    write_raw(dir + "/distNN.bin", dist.data(), dist.size() * sizeof(double));
}

// Rebuild a handle from a saved structure (the reference's file set) and the
// points, then compress and pack it.  The blocked pair arrays, level sets and
// coarsened partitions are taken from the files as they are, so D and B are
// packed in the saved pair order.
static Handle *load_bin(const std::string &dir, const double *X, int n, int d, const mxh_config &cfg) {
    std::unique_ptr<Handle> H(new Handle());
    H->cfg = cfg;
    validate(cfg, n, d);
    H->n = n;
    H->d = d;
    H->X.assign(X, X + (size_t)n * d);
    H->f64["X"] = H->X;
    auto R = [&](const char *name, std::vector<int32_t> &v, bool required) {
        bool ok = read_raw(dir + "/" + name, v);
        if (!ok && required) throw std::runtime_error(std::string("missing ") + name);
        return ok;
    };
    std::vector<int32_t> dim, lm, lids, lo, ll, lc, rc, nx, ny, fx, fy, nbs, nbk, fbs, fbk, cl, wp, cidx, lvl, idx,
        nnidx;
    R("dim.bin", dim, true);
    R("leafmap.bin", lm, true);
    R("lids.bin", lids, true);
    R("lidsoffset.bin", lo, true);
    R("lidslen.bin", ll, true);
    R("lchildren.bin", lc, true);
    R("rchildren.bin", rc, true);
    R("nxval.bin", nx, true);
    R("nyval.bin", ny, true);
    R("fxval.bin", fx, true);
    R("fyval.bin", fy, true);
    bool blocked = R("nblockset.bin", nbs, false);
    if (blocked) {
        R("nblocks.bin", nbk, true);
        R("fblockset.bin", fbs, true);
        R("fblocks.bin", fbk, true);
    }
    R("clevelset.bin", cl, true);
    R("wpart.bin", wp, true);
    R("cidx.bin", cidx, true);
    R("levelset.bin", lvl, true);
    R("idx.bin", idx, true);
    R("NN.bin", nnidx, true);
    std::vector<double> dist;
    if (!read_raw(dir + "/distNN.bin", dist)) throw std::runtime_error("missing distNN.bin");

    const int L = (int)dim.size();
    int depth = 0;
    while ((1 << depth) < L) ++depth;
    if (L < 2 || (1 << depth) != L) throw std::runtime_error("leaf count must be a power of two (perfect tree)");
    const int nn = 2 * L - 1, f = L - 1;
    if ((int)lc.size() != nn || (int)rc.size() != nn || (int)lm.size() != nn || (int)lo.size() != nn ||
        (int)ll.size() != nn)
        throw std::runtime_error("per-node arrays must have 2*nleaf-1 entries");
    for (int v = 0; v < nn; ++v) {
        bool leaf = v >= f;
        if (leaf ? (lc[v] != -1 || rc[v] != -1 || lm[v] != v - f) : (lc[v] != 2 * v + 1 || rc[v] != 2 * v + 2))
            throw std::runtime_error("tree is not a perfect binary tree in BFS numbering");
    }
    H->depth = depth;
    H->nleaf = L;
    H->numnodes = nn;
    H->pts.assign(nn, {});
    std::vector<char> seen(n, 0);
    for (int v = 1; v < nn; ++v) {
        if (lo[v] < 0 || ll[v] < 0 || (size_t)lo[v] + ll[v] > lids.size())
            throw std::runtime_error("lids offsets out of range");
        H->pts[v].assign(lids.begin() + lo[v], lids.begin() + lo[v] + ll[v]);
        for (int p : H->pts[v])
            if (p < 0 || p >= n) throw std::runtime_error("point index out of range");
    }
    for (int l = 0; l < L; ++l) {
        if ((int)H->pts[f + l].size() != dim[l]) throw std::runtime_error("dim.bin disagrees with lidslen");
        for (int p : H->pts[f + l]) {
            if (seen[p]) throw std::runtime_error("point in two leaves");
            seen[p] = 1;
        }
    }
    for (int p = 0; p < n; ++p)
        if (!seen[p]) throw std::runtime_error("lids is not a permutation of the points");
    H->pts[0].resize(n);
    for (int p = 0; p < n; ++p) H->pts[0][p] = p;
    H->sflevel.assign(depth + 1, {});
    for (int l = 0; l <= depth; ++l) {
        int dl = depth - l;
        for (int v = (1 << dl) - 1; v < (1 << (dl + 1)) - 1; ++v) H->sflevel[l].push_back(v);
    }
    if (nx.size() != ny.size() || fx.size() != fy.size()) throw std::runtime_error("pair arrays differ in length");
    H->nearn.assign(nn, {});
    H->farn.assign(nn, {});
    for (size_t k = 0; k < nx.size(); ++k) {
        if (nx[k] < 0 || nx[k] >= L || ny[k] < 0 || ny[k] >= L) throw std::runtime_error("near pair out of range");
        H->nearn[f + nx[k]].push_back(f + ny[k]);
    }
    for (size_t k = 0; k < fx.size(); ++k) {
        if (fx[k] <= 0 || fx[k] >= nn || fy[k] <= 0 || fy[k] >= nn) throw std::runtime_error("far pair out of range");
        H->farn[fx[k]].push_back(fy[k]);
    }
    for (int v = 0; v < nn; ++v) {
        std::sort(H->nearn[v].begin(), H->nearn[v].end());
        std::sort(H->farn[v].begin(), H->farn[v].end());
    }
    H->is_hss = true;
    for (int v = f; v < nn; ++v)
        if (H->nearn[v].size() != 1) H->is_hss = false;
    // coarsened partitions from clevelset / wpart / cidx
    if (cl.empty() || wp.empty() || (int)cidx.size() != nn - 1) throw std::runtime_error("bad coarsening arrays");
    H->postw.assign(cl.size() - 1, {});
    for (size_t g = 0; g + 1 < cl.size(); ++g)
        for (int p = cl[g]; p < cl[g + 1]; ++p) {
            if (p < 0 || p + 1 >= (int)wp.size()) throw std::runtime_error("clevelset out of range");
            H->postw[g].push_back({});
            for (int j = wp[p]; j < wp[p + 1]; ++j) H->postw[g].back().push_back(cidx[j]);
        }
    // neighbour table
    if (nnidx.size() != dist.size() || nnidx.empty() || nnidx.size() % (size_t)n)
        throw std::runtime_error("NN.bin / distNN.bin do not hold n*k entries");
    H->knn_k = (int)(nnidx.size() / n);
    H->NN.resize(nnidx.size());
    for (size_t i = 0; i < nnidx.size(); ++i) H->NN[i] = {dist[i], nnidx[i]};
    int saved = omp_get_max_threads();
    if (cfg.nthreads > 0) omp_set_num_threads(cfg.nthreads);
    try {
        emit_structure(*H);
        auto &I = H->i32;
        I["levelset"] = lvl;
        I["idx_level"] = idx;
        I["clevelset"] = cl;
        I["wpart"] = wp;
        I["idx"] = cidx;
        if (blocked && !H->is_hss) {
            I["nblockSet"] = nbs;
            I["nblocks"] = nbk;
            I["npairx"] = nx;
            I["npairy"] = ny;
            I["fblockSet"] = fbs;
            I["fblocks"] = fbk;
            I["fpairx"] = fx;
            I["fpairy"] = fy;
            H->scal["nl"] = (int64_t)nbs.size() - 1;
            H->scal["fl"] = (int64_t)fbs.size() - 1;
        }
        compress(*H);
        balance_partitions(*H);
        emit_cds(*H);
    } catch (...) {
        omp_set_num_threads(saved);
        throw;
    }
    omp_set_num_threads(saved);
    return H.release();
}

}  // namespace mxh

using mxh::Handle;

struct mxh_handle : public Handle {};

template <class T>
static int get_arr(const std::map<std::string, std::vector<T>> &m, const char *name, const T **p,
                   int64_t *len) {
    auto it = m.find(name);
    if (it == m.end()) {
        mxh::g_last_error = std::string("unknown array ") + name;
        return -2;
    }
    *p = it->second.data();
    *len = (int64_t)it->second.size();
    return 0;
}

extern "C" {

void mxh_config_default(mxh_config *c) {
    c->leaf_size = 256;
    c->max_rank = 256;
    c->tol = 1e-5;
    c->budget = 0.03;
    c->h = 5.0;
    c->kernel = MXH_KERNEL_GAUSSIAN;
    c->knn_k = 32;
    c->near_block = 2;
    c->far_block = 4;
    c->agg_levels = 2;
    c->bottom_levels = 2;
    c->nbins = 12;
    c->nthreads = 0;
    c->seed = 0x4D617452ULL;
    c->skip_blocks = 0;
    c->all_rows = 0;
}

static mxh_handle *build_c(const double *X, int32_t n, int32_t d, const mxh_config *cfg, bool full) {
    if (!X || !cfg) {
        mxh::g_last_error = "null argument";
        return nullptr;
    }
    try {
        Handle *h = mxh::build(X, n, d, cfg, full);
        mxh_handle *out = new mxh_handle();
        static_cast<Handle &>(*out) = std::move(*h);
        delete h;
        return out;
    } catch (const std::exception &e) {
        mxh::g_last_error = e.what();
        return nullptr;
    }
}

mxh_handle *mxh_build(const double *X, int32_t n, int32_t d, const mxh_config *cfg) {
    return build_c(X, n, d, cfg, true);
}

mxh_handle *mxh_build_structure(const double *X, int32_t n, int32_t d, const mxh_config *cfg) {
    return build_c(X, n, d, cfg, false);
}

void mxh_free(mxh_handle *h) { delete h; }

int64_t mxh_get_scalar(const mxh_handle *h, const char *name) {
    if (!h || !name) return -1;
    auto it = h->scal.find(name);
    return it == h->scal.end() ? -1 : it->second;
}

int mxh_get_i32(const mxh_handle *h, const char *name, const int32_t **p, int64_t *len) {
    if (!h || !name || !p || !len) return -1;
    return get_arr(h->i32, name, p, len);
}
int mxh_get_u64(const mxh_handle *h, const char *name, const uint64_t **p, int64_t *len) {
    if (!h || !name || !p || !len) return -1;
    return get_arr(h->u64, name, p, len);
}
int mxh_get_f64(const mxh_handle *h, const char *name, const double **p, int64_t *len) {
    if (!h || !name || !p || !len) return -1;
    return get_arr(h->f64, name, p, len);
}

// transform (matroxTest-V1/hcds.h:463-485): TW[wrow[l]*nrhs + j*Dim_l + k] = W[j*n + lids_l[k]]
int mxh_transform(const mxh_handle *h, const double *W, int32_t nrhs, double *TW) {
    if (!h || !W || !TW || nrhs < 0) return -1;
    const int f = h->leaf_first(), n = h->n;
    const auto &wr = h->i32.at("wrow");
#pragma omp parallel for schedule(dynamic)
    for (int l = 0; l < h->nleaf; ++l) {
        const auto &p = h->pts[f + l];
        const int dim = (int)p.size();
        double *tw = TW + (size_t)wr[l] * nrhs;
        for (int j = 0; j < nrhs; ++j)
            for (int k = 0; k < dim; ++k) tw[(size_t)j * dim + k] = W[(size_t)j * n + p[k]];
    }
    return 0;
}

int mxh_untransform(const mxh_handle *h, const double *TW, int32_t nrhs, double *W) {
    if (!h || !W || !TW || nrhs < 0) return -1;
    const int f = h->leaf_first(), n = h->n;
    const auto &wr = h->i32.at("wrow");
#pragma omp parallel for schedule(dynamic)
    for (int l = 0; l < h->nleaf; ++l) {
        const auto &p = h->pts[f + l];
        const int dim = (int)p.size();
        const double *tw = TW + (size_t)wr[l] * nrhs;
        for (int j = 0; j < nrhs; ++j)
            for (int k = 0; k < dim; ++k) W[(size_t)j * n + p[k]] = tw[(size_t)j * dim + k];
    }
    return 0;
}

int mxh_gaussian_mixture(double *X, int64_t n, int32_t d, int32_t ncomp, double mean_range,
                         double sig_lo, double sig_hi, int32_t diag, int32_t standardise,
                         uint64_t seed) {
    if (!X || n < 1 || d < 1 || ncomp < 1) return -1;
    mxh::SplitMix64 g(mxh::stream_seed(seed, 0xC0FFEE));
    std::vector<double> mu((size_t)ncomp * d), sg((size_t)ncomp * d);
    for (int c = 0; c < ncomp; ++c) {
        for (int p = 0; p < d; ++p) mu[(size_t)c * d + p] = (2 * g.uniform() - 1) * mean_range;
        double s_iso = sig_lo + (sig_hi - sig_lo) * g.uniform();
        for (int p = 0; p < d; ++p)
            sg[(size_t)c * d + p] = diag ? sig_lo + (sig_hi - sig_lo) * g.uniform() : s_iso;
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        mxh::SplitMix64 gi(mxh::stream_seed(seed, 0xA11CE, (uint64_t)i));
        int c = (int)gi.below((uint64_t)ncomp);
        uint64_t ps = mxh::stream_seed(seed, 0xB0B, (uint64_t)i);
        for (int p = 0; p < d; p += 2) {
            double z0, z1;
            mxh::normal_pair(ps, (uint64_t)p / 2, &z0, &z1);
            X[i * d + p] = mu[(size_t)c * d + p] + sg[(size_t)c * d + p] * z0;
            if (p + 1 < d) X[i * d + p + 1] = mu[(size_t)c * d + p + 1] + sg[(size_t)c * d + p + 1] * z1;
        }
    }
    if (standardise) {
        // fixed chunks summed in a fixed order: bit-identical for any thread count
        const int nch = 64;
        for (int p = 0; p < d; ++p) {
            double ps[nch], ps2[nch];
#pragma omp parallel for schedule(static)
            for (int c = 0; c < nch; ++c) {
                int64_t b = n * c / nch, e = n * (c + 1) / nch;
                double a = 0, a2 = 0;
                for (int64_t i = b; i < e; ++i) {
                    double x = X[i * d + p];
                    a += x;
                    a2 += x * x;
                }
                ps[c] = a;
                ps2[c] = a2;
            }
            double s = 0, s2 = 0;
            for (int c = 0; c < nch; ++c) {
                s += ps[c];
                s2 += ps2[c];
            }
            double mean = s / n;
            double var = s2 / n - mean * mean;
            double sd = var > 0 ? std::sqrt(var) : 1.0;
#pragma omp parallel for
            for (int64_t i = 0; i < n; ++i) X[i * d + p] = (X[i * d + p] - mean) / sd;
        }
    }
    return 0;
}

int mxh_normal(double *out, int64_t count, uint64_t seed) {
    if (!out || count < 0) return -1;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (count + 1) / 2; ++i) {
        double z0, z1;
        mxh::normal_pair(seed, (uint64_t)i, &z0, &z1);
        out[2 * i] = z0;
        if (2 * i + 1 < count) out[2 * i + 1] = z1;
    }
    return 0;
}

int mxh_normal_f32(float *out, int64_t count, uint64_t seed) {
    if (!out || count < 0) return -1;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (count + 1) / 2; ++i) {
        double z0, z1;
        mxh::normal_pair(seed, (uint64_t)i, &z0, &z1);
        out[2 * i] = (float)z0;
        if (2 * i + 1 < count) out[2 * i + 1] = (float)z1;
    }
    return 0;
}

int mxh_save_bin(const mxh_handle *h, const char *dir) {
    if (!h || !dir) return -1;
    try {
        mxh::save_bin(*h, dir);
        return 0;
    } catch (const std::exception &e) {
        mxh::g_last_error = e.what();
        return -3;
    }
}

mxh_handle *mxh_load_bin(const char *dir, const double *X, int32_t n, int32_t d, const mxh_config *cfg) {
    if (!dir || !X || !cfg) {
        mxh::g_last_error = "null argument";
        return nullptr;
    }
    try {
        Handle *h = mxh::load_bin(dir, X, n, d, *cfg);
        mxh_handle *out = new mxh_handle();
        static_cast<Handle &>(*out) = std::move(*h);
        delete h;
        return out;
    } catch (const std::exception &e) {
        mxh::g_last_error = e.what();
        return nullptr;
    }
}

const char *mxh_last_error(void) { return mxh::g_last_error.c_str(); }

}  // extern "C"
