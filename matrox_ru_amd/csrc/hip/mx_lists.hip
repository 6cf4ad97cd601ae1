// Structure checks and the per-destination task lists of an evaluation
// (forward and adjoint, two coupling schedules, two kernel kinds) and of the
// block gradients.
#include "mx_internal.h"

#include <set>

namespace mx {

void compute_launch_flops(TaskLists &L) {
    L.flops.assign(L.launches.size(), 0.0);
    for (int &w : L.last_write) w = -1;
    std::vector<int> seen(L.tasks.size(), -1);
    for (size_t li = 0; li < L.launches.size(); ++li) {
        for (int t = L.launches[li].tile_offset; t < L.launches[li].tile_offset + L.launches[li].ntiles; ++t) {
            int task = L.tiles[t].x;
            if (seen[task] == (int)li) continue;   // count each task once per launch
            seen[task] = (int)li;
            const Task &tk = L.tasks[task];
            if (tk.seg_end > tk.seg_begin) L.last_write[tk.c_buf] = (int)li;
            for (int sg = tk.seg_begin; sg < tk.seg_end; ++sg) L.flops[li] += 2.0 * tk.m * (double)L.segs[sg].k;
        }
    }
}

Topology topology(const mx_structure &s) {
    Topology t;
    const int nn = s.numnodes;
    t.parent.assign(nn, -1);
    t.depth.assign(nn, -1);
    t.node_of_leaf.assign(s.nleaf, -1);
    for (int v = 0; v < nn; ++v) {
        int l = s.lchildren[v], r = s.rchildren[v];
        arg_check((l == -1) == (r == -1), "node " + std::to_string(v) + " has exactly one child");
        if (l != -1) {
            arg_check(l > 0 && l < nn && r > 0 && r < nn && l != r, "child index out of range");
            arg_check(t.parent[l] == -1 && t.parent[r] == -1, "node with two parents");
            t.parent[l] = v;
            t.parent[r] = v;
        } else {
            int lf = s.lm[v];
            arg_check(lf >= 0 && lf < s.nleaf, "leaf node " + std::to_string(v) + " has no leaf index");
            arg_check(t.node_of_leaf[lf] == -1, "leaf index used twice");
            t.node_of_leaf[lf] = v;
        }
    }
    arg_check(t.parent[0] == -1, "node 0 must be the root");
    // depth by BFS from the root
    std::vector<int> q{0};
    t.depth[0] = 0;
    for (size_t i = 0; i < q.size(); ++i) {
        int v = q[i];
        if (s.lchildren[v] != -1) {
            for (int c : {s.lchildren[v], s.rchildren[v]}) {
                t.depth[c] = t.depth[v] + 1;
                q.push_back(c);
            }
        }
    }
    arg_check((int)q.size() == nn, "tree does not reach every node from the root");
    int maxd = 0;
    for (int v = 0; v < nn; ++v) maxd = std::max(maxd, t.depth[v]);
    t.levels.assign(maxd + 1, {});
    for (int v = 1; v < nn; ++v) t.levels[t.depth[v]].push_back(v);
    for (int l = 0; l < s.nleaf; ++l) arg_check(t.node_of_leaf[l] != -1, "leaf index without node");
    return t;
}

static int cols_of(const mx_structure &s, int v) {
    if (s.lchildren[v] == -1) return s.Ddim[s.lm[v]];
    return s.slen[s.lchildren[v]] + s.slen[s.rchildren[v]];
}

void validate(const mx_structure &s, const Topology &t) {
    const int nn = s.numnodes;
    for (int l = 0; l < s.nleaf; ++l) arg_check(s.Ddim[l] >= 0, "negative leaf size");
    for (int v = 0; v < nn; ++v) arg_check(s.slen[v] >= 0, "negative skeleton length");
    for (int v = 1; v < nn; ++v) {
        int64_t need = (int64_t)s.VTptr[v] + (int64_t)s.slen[v] * cols_of(s, v);
        arg_check(s.slen[v] == 0 || (s.VTptr[v] >= 0 && need <= s.VT_len),
                  "VT block of node " + std::to_string(v) + " exceeds VT");
    }
    for (int64_t k = 0; k < s.ncount; ++k) {
        int x = s.is_hss ? (int)k : s.npairx[k], y = s.is_hss ? (int)k : s.npairy[k];
        arg_check(x >= 0 && x < s.nleaf && y >= 0 && y < s.nleaf, "near pair leaf out of range");
        int64_t need = (int64_t)s.Dptr[k] + (int64_t)s.Ddim[x] * s.Ddim[y];
        arg_check(need <= s.D_len, "near block " + std::to_string(k) + " exceeds D");
    }
    for (int64_t k = 0; k < s.fcount; ++k) {
        int x, y;
        if (s.is_hss) {
            x = (int)k + 1;
            arg_check(x < nn, "HSS far count exceeds numnodes-1");
            int p = t.parent[x];
            y = s.lchildren[p] == x ? s.rchildren[p] : s.lchildren[p];
        } else {
            x = s.fpairx[k];
            y = s.fpairy[k];
        }
        arg_check(x > 0 && x < nn && y > 0 && y < nn, "far pair node out of range");
        int64_t need = (int64_t)s.Bptr[k] + (int64_t)s.slen[x] * s.slen[y];
        arg_check(need <= s.B_len, "far block " + std::to_string(k) + " exceeds B");
    }
    for (int l = 0; l < s.nleaf; ++l) arg_check(s.wrow[l] >= 0, "negative leaf row offset");
    for (int v = 0; v < nn; ++v) arg_check(s.skelrow[v] >= 0, "negative skeleton row offset");
}

static int add_tiles(TaskLists &L, int task, int m) {
    const int bm = L.task_kind[task] == 1 ? X3_BM : SG_BM;
    int n = (m + bm - 1) / bm;
    for (int i = 0; i < n; ++i) L.tiles.push_back(make_int2(task, i));
    return n;
}

// Order of the tiles of every launch.  xcd_remap hands XCD x the x-th
// contiguous eighth of a launch's work items (the column tiles of a tile are
// adjacent), and an XCD runs about 64 work items at a time: at nrhs = 1024
// the row and column tiles of four 256-row tasks.  Its 4 MB L2 holds little
// more than one K-chunk of their operands, so a B operand (a W or skeleton
// panel, read by every task whose list names it) is served from L2 to a
// second task only when both read it at about the same time.
//
// The BF16x3 kernel's launches of at least 64 tasks (the big ones) are
// therefore ordered in groups of four tasks that share B operands: greedily,
// the task with the most segments, then repeatedly the task sharing the most
// operands with the group.  Within a group the segments are reordered so
// that a shared operand sits at the same position in every task of the
// group that reads it (slot by slot, the operand the most unplaced tasks
// still need).  Groups go to eight bins of about equal work (most work
// first, to the least loaded bin), each bin runs its groups longest task
// first (long lists start early), and the bins are padded with empty tiles
// to the size of the largest, so that bin x is exactly the eighth xcd_remap
// gives XCD x.
//
// Smaller launches, and the FP64 / FP32 kernel's (64 x 64 tiles: a window of
// 64 work items holds about two 256-row tasks, half a group; the grouping
// measured 2-4% slower there), keep the tiles in structural order (tree
// order: neighbouring nodes share skeleton blocks), cut into eight
// contiguous groups of about equal work, each sorted longest first and
// padded the same way; this spreads a few tasks' tiles over all eight XCDs.
static int64_t seg_key(const Seg &g) { return ((int64_t)g.b_buf << 48) ^ g.b_row; }

// Slot by slot, the B operand that the most tasks of the group still needing
// a segment at this slot share (ties: the smallest key) goes next in each of
// them.  Tasks are bits of a mask (a group has at most 32).
static void reorder_group_segments(TaskLists &L, const std::vector<int> &tasks) {
    const int G = (int)tasks.size();
    std::vector<int64_t> keys;
    for (int t : tasks)
        for (int sg = L.tasks[t].seg_begin; sg < L.tasks[t].seg_end; ++sg) keys.push_back(seg_key(L.segs[sg]));
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    const int K = (int)keys.size();
    // per task: its segments by key index (ascending), a cursor per key
    std::vector<std::vector<std::pair<int, int>>> by(G);   // (key index, segment)
    std::vector<uint32_t> holders(K, 0);
    std::vector<int> left(G);
    for (int j = 0; j < G; ++j) {
        const Task &t = L.tasks[tasks[j]];
        for (int sg = t.seg_begin; sg < t.seg_end; ++sg) {
            const int ki = (int)(std::lower_bound(keys.begin(), keys.end(), seg_key(L.segs[sg])) - keys.begin());
            by[j].push_back({ki, sg});
            holders[ki] |= 1u << j;
        }
        std::stable_sort(by[j].begin(), by[j].end(),
                         [](const std::pair<int, int> &x, const std::pair<int, int> &y) { return x.first < y.first; });
        left[j] = (int)by[j].size();
    }
    std::vector<std::vector<char>> used(G);
    for (int j = 0; j < G; ++j) used[j].assign(by[j].size(), 0);
    std::vector<std::vector<Seg>> seq(G);
    auto take = [&](int j, int ki) {   // the first unused segment of task j with key ki
        auto it = std::lower_bound(by[j].begin(), by[j].end(), std::make_pair(ki, -1));
        size_t i = (size_t)(it - by[j].begin());
        while (used[j][i]) ++i;
        used[j][i] = 1;
        seq[j].push_back(L.segs[by[j][i].second]);
        --left[j];
        if (i + 1 >= by[j].size() || by[j][i + 1].first != ki || used[j][i + 1]) holders[ki] &= ~(1u << j);
    };
    for (;;) {
        uint32_t open = 0;
        for (int j = 0; j < G; ++j)
            if (left[j] > 0) open |= 1u << j;
        if (!open) break;
        while (open) {   // one position of every open task
            int best = -1, bc = 0;
            for (int ki = 0; ki < K; ++ki) {
                const int c = __builtin_popcount(holders[ki] & open);
                if (c > bc) {
                    bc = c;
                    best = ki;
                }
            }
            const uint32_t placed = holders[best] & open;
            for (int j = 0; j < G; ++j)
                if (placed >> j & 1u) take(j, best);
            open &= ~placed;
        }
    }
    for (int j = 0; j < G; ++j) std::copy(seq[j].begin(), seq[j].end(), L.segs.begin() + L.tasks[tasks[j]].seg_begin);
}

static void order_tiles(TaskLists &L, int bk, bool group_by_operand) {
    if (L.dummy < 0) {
        Task tk{};
        tk.c_buf = VEC_USKEL;
        L.dummy = (int)L.tasks.size();
        L.tasks.push_back(tk);
        L.task_kind.push_back(L.task_kind.empty() ? 0 : L.task_kind.back());
    }
    const int ng = 8, G = 4, min_tasks = 64;
    auto tile_cost = [&](int task) {
        const Task &tk = L.tasks[task];
        int64_t cost = 2;   // prologue and epilogue
        for (int sg = tk.seg_begin; sg < tk.seg_end; ++sg) cost += (L.segs[sg].k + bk - 1) / bk;
        return cost;
    };
    std::vector<int2> out;
    std::vector<Launch> lns;
    std::vector<int> local(L.tasks.size(), -1);
    for (const Launch &ln : L.launches) {
        const int T = ln.ntiles;
        // the launch's tasks in structural order, with their row tiles
        std::vector<int> tk;
        std::vector<std::vector<int>> rts;
        for (int i = 0; i < T; ++i) {
            const int2 tl = L.tiles[ln.tile_offset + i];
            if (local[tl.x] < 0) {
                local[tl.x] = (int)tk.size();
                tk.push_back(tl.x);
                rts.emplace_back();
            }
            rts[local[tl.x]].push_back(tl.y);
        }
        for (int t : tk) local[t] = -1;
        const int n = (int)tk.size();
        std::vector<std::vector<int2>> bins(ng);
        if (n < min_tasks || !group_by_operand) {
            std::vector<std::pair<int64_t, int2>> v;
            v.reserve(T);
            int64_t total = 0;
            for (int i = 0; i < T; ++i) {
                const int2 tl = L.tiles[ln.tile_offset + i];
                v.push_back({tile_cost(tl.x), tl});
                total += v.back().first;
            }
            std::vector<int> cut(ng + 1, T);
            cut[0] = 0;
            int64_t acc = 0;
            int g = 1;
            for (int i = 0; i < T && g < ng; ++i) {
                acc += v[i].first;
                while (g < ng && acc * ng >= total * g) cut[g++] = i + 1;
            }
            for (int q = 0; q < ng; ++q) {
                std::stable_sort(v.begin() + cut[q], v.begin() + cut[q + 1],
                                 [](const std::pair<int64_t, int2> &a, const std::pair<int64_t, int2> &b) {
                                     return a.first > b.first;
                                 });
                for (int i = cut[q]; i < cut[q + 1]; ++i) bins[q].push_back(v[i].second);
            }
        } else {
            std::vector<std::vector<int>> src(n);   // distinct B operands (dense indices) of each task
            std::vector<int64_t> all;
            std::vector<int64_t> cost(n);
            for (int i = 0; i < n; ++i) {
                const Task &t = L.tasks[tk[i]];
                for (int sg = t.seg_begin; sg < t.seg_end; ++sg) all.push_back(seg_key(L.segs[sg]));
                cost[i] = tile_cost(tk[i]);
            }
            std::sort(all.begin(), all.end());
            all.erase(std::unique(all.begin(), all.end()), all.end());
            std::vector<std::vector<int>> users(all.size());
            for (int i = 0; i < n; ++i) {
                const Task &t = L.tasks[tk[i]];
                for (int sg = t.seg_begin; sg < t.seg_end; ++sg)
                    src[i].push_back((int)(std::lower_bound(all.begin(), all.end(), seg_key(L.segs[sg])) - all.begin()));
                std::sort(src[i].begin(), src[i].end());
                src[i].erase(std::unique(src[i].begin(), src[i].end()), src[i].end());
                for (int k : src[i]) users[k].push_back(i);
            }
            std::vector<int> seeds(n);
            for (int i = 0; i < n; ++i) seeds[i] = i;
            std::stable_sort(seeds.begin(), seeds.end(), [&](int a, int b) { return src[a].size() > src[b].size(); });
            std::vector<char> used(n, 0), inU(all.size(), 0);
            std::vector<int> score(n, 0);
            std::vector<std::vector<int>> groups;
            size_t sp = 0;
            int lowest = 0;
            for (;;) {
                while (sp < seeds.size() && used[seeds[sp]]) ++sp;
                if (sp == seeds.size()) break;
                std::vector<int> g{seeds[sp]};
                used[seeds[sp]] = 1;
                std::vector<int> U;   // operands of the group
                auto add = [&](int task) {
                    for (int k : src[task])
                        if (!inU[k]) {
                            inU[k] = 1;
                            U.push_back(k);
                        }
                };
                add(g[0]);
                while ((int)g.size() < G) {
                    std::vector<int> touched;
                    for (int k : U)
                        for (int c : users[k])
                            if (!used[c] && score[c]++ == 0) touched.push_back(c);
                    int best = -1;
                    for (int c : touched)
                        if (best < 0 || score[c] > score[best] || (score[c] == score[best] && c < best)) best = c;
                    for (int c : touched) score[c] = 0;
                    if (best < 0) {   // nothing shared: the next task in structural order
                        while (lowest < n && used[lowest]) ++lowest;
                        if (lowest == n) break;
                        best = lowest;
                    }
                    used[best] = 1;
                    g.push_back(best);
                    add(best);
                }
                for (int k : U) inU[k] = 0;
                groups.push_back(g);
            }
            const int ngr = (int)groups.size();
            std::vector<int64_t> work(ngr, 0), longest(ngr, 0);
            for (int q = 0; q < ngr; ++q) {
                std::vector<int> ids;
                for (int j : groups[q]) {
                    work[q] += cost[j] * (int64_t)rts[j].size();
                    longest[q] = std::max(longest[q], cost[j]);
                    ids.push_back(tk[j]);
                }
                reorder_group_segments(L, ids);
            }
            std::vector<int> order(ngr);
            for (int q = 0; q < ngr; ++q) order[q] = q;
            std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work[a] > work[b]; });
            std::vector<int64_t> load(ng, 0);
            std::vector<std::vector<int>> bg(ng);
            for (int q : order) {
                const int b = (int)(std::min_element(load.begin(), load.end()) - load.begin());
                bg[b].push_back(q);
                load[b] += work[q];
            }
            for (int b = 0; b < ng; ++b) {
                std::stable_sort(bg[b].begin(), bg[b].end(), [&](int a, int c) { return longest[a] > longest[c]; });
                for (int q : bg[b])
                    for (int j : groups[q])
                        for (int r : rts[j]) bins[b].push_back(make_int2(tk[j], r));
            }
        }
        size_t gmax = 0;
        for (const auto &b : bins) gmax = std::max(gmax, b.size());
        Launch nl{(int)out.size(), (int)gmax * ng};
        for (const auto &b : bins) {
            out.insert(out.end(), b.begin(), b.end());
            for (size_t i = b.size(); i < gmax; ++i) out.push_back(make_int2(L.dummy, 0));
        }
        lns.push_back(nl);
    }
    L.tiles.swap(out);
    L.launches.swap(lns);
}

// The coarse schedule: runs of consecutive launches of the fused lists merge
// into one launch.  A segment A_c x_P of a launch, over a skeleton block
// x_P = sum_s A_s B_s written by producer task P (beta = 0, its only writer)
// in an earlier launch of the same run, becomes the segments (A_c A_s) B_s,
// one per segment of P, with the composite A_c A_s formed in FP64 when the
// plan is built.  P's own segments are already over blocks written before
// the run, so a composite's second factor is a stored block or an earlier
// composite (chains as long as the run).  Every skeleton block is still
// written (the block gradients read them all).  Launch j+1 joins the run
// of launch j when that adds at most max_extra flops per right-hand side
// column (2 m (sum_s k_s - k_c) per rewritten segment), leaves every task
// within maxseg segments, and does not lengthen the run's longest task by
// more than launch j+1's own longest task plus launch_chunks (the top of the
// tree is latency-bound: a launch lasts about as long as its longest task
// plus a fixed cost), and none of its tasks writes a block the run writes
// or reads (the split lists' per-level transfers accumulate onto the
// couplings' blocks); otherwise it starts a run of its own.  The UP launches
// of the lowest two levels merge for nothing when the leaf blocks are as
// tall as the ranks, the top-of-tree launches for little.
static void coarsen(TaskLists &L, double max_extra, int maxseg, int bk, int launch_chunks) {
    const int n = (int)L.launches.size();
    std::vector<int> launch_of(L.tasks.size(), -1);
    std::vector<std::vector<int>> tasks_of(n);
    for (int li = 0; li < n; ++li)
        for (int t = L.launches[li].tile_offset; t < L.launches[li].tile_offset + L.launches[li].ntiles; ++t) {
            const int task = L.tiles[t].x;
            if (launch_of[task] < 0) {
                launch_of[task] = li;
                tasks_of[li].push_back(task);
            }
        }
    // the producer of every skeleton block (-1: several writers)
    std::map<std::pair<int, int64_t>, int> prod;
    for (size_t t = 0; t < L.tasks.size(); ++t) {
        const Task &tk = L.tasks[t];
        if (launch_of[t] < 0 || (tk.c_buf != VEC_WSKEL && tk.c_buf != VEC_USKEL)) continue;
        auto key = std::make_pair((int)tk.c_buf, tk.c_row);
        auto it = prod.find(key);
        if (it == prod.end()) prod[key] = (int)t;
        else it->second = -1;
    }
    // every task's segments as they will be (over blocks written before its run)
    std::vector<std::vector<Seg>> segs_of(L.tasks.size());
    for (size_t t = 0; t < L.tasks.size(); ++t)
        segs_of[t].assign(L.segs.begin() + L.tasks[t].seg_begin, L.segs.begin() + L.tasks[t].seg_end);
    auto composite_of = [&](const Seg &outer, const Seg &inner, int m) {
        Composite c{};
        c.a1_off = outer.a_off;
        c.a1_buf = outer.a_buf;
        c.a1_trans = outer.a_trans;
        c.a1_lda = outer.lda;
        c.a2_off = inner.a_off;   // a stored block, or (MAT_COMP) an earlier composite
        c.a2_buf = inner.a_buf;
        c.a2_trans = inner.a_trans;
        c.a2_lda = inner.lda;
        c.m = m;
        c.kmid = outer.k;
        c.n = inner.k;
        return c;
    };
    // a segment of a joining launch: kept (P < 0), or the composite of it
    // with segment q of its producer P
    struct Piece {
        Seg sg;
        int P, q;
    };
    // a task's length in K chunks: a launch of latency-bound tasks lasts about
    // as long as its longest task, plus a fixed cost of some chunks' worth
    auto chunks = [&](const std::vector<Seg> &sg) {
        int c = 0;
        for (const Seg &x : sg) c += (x.k + bk - 1) / bk;
        return c;
    };
    std::vector<char> joins(n, 0);   // launch joins the run of the launch before it
    int run0 = 0;                    // first launch of the current run
    int run_len = 0;                 // longest task of the current run, in chunks
    // blocks the run's tasks write and read: a launch whose tasks write one
    // of them cannot join (its tasks would run beside the run's, in any order)
    std::set<std::pair<int, int64_t>> run_io;
    auto note_run = [&](int li) {
        for (int t : tasks_of[li]) {
            run_io.insert(std::make_pair((int)L.tasks[t].c_buf, L.tasks[t].c_row));
            for (const Seg &sg : segs_of[t]) run_io.insert(std::make_pair((int)sg.b_buf, sg.b_row));
        }
    };
    for (int t : tasks_of[0]) run_len = std::max(run_len, chunks(segs_of[t]));
    note_run(0);
    for (int j = 1; j < n; ++j) {
        double extra = 0.0;
        bool ok = true;
        int own_len = 0, new_len = 0;   // launch j's longest task before and after rewriting
        std::vector<std::pair<int, std::vector<Piece>>> rewritten;
        for (int t : tasks_of[j])
            if (run_io.count(std::make_pair((int)L.tasks[t].c_buf, L.tasks[t].c_row))) ok = false;
        for (int t : tasks_of[j]) {
            if (!ok) break;
            const Task &tk = L.tasks[t];
            std::vector<Piece> out;
            for (const Seg &sg : segs_of[t]) {
                int P = -1;
                if (sg.b_buf == VEC_WSKEL || sg.b_buf == VEC_USKEL) {
                    auto it = prod.find(std::make_pair((int)sg.b_buf, sg.b_row));
                    if (it != prod.end()) {
                        if (it->second < 0) ok = false;   // a block with several writers
                        else if (launch_of[it->second] >= run0 && launch_of[it->second] < j) P = it->second;
                    }
                }
                if (P < 0) {
                    out.push_back(Piece{sg, -1, -1});
                    continue;
                }
                const Task &pt = L.tasks[P];
                if (pt.beta != 0 || pt.m != sg.k) {
                    ok = false;
                    break;
                }
                double kp = 0.0;
                for (size_t q = 0; q < segs_of[P].size(); ++q) {
                    out.push_back(Piece{sg, P, (int)q});
                    kp += segs_of[P][q].k;
                }
                extra += 2.0 * tk.m * (kp - sg.k);
            }
            if (!ok || (int)out.size() > maxseg) {
                ok = false;
                break;
            }
            own_len = std::max(own_len, chunks(segs_of[t]));
            int c = 0;
            for (const Piece &pc : out) c += ((pc.P < 0 ? pc.sg.k : segs_of[pc.P][pc.q].k) + bk - 1) / bk;
            new_len = std::max(new_len, c);
            rewritten.push_back({t, std::move(out)});
        }
        // the merged launch must not outlast the two it replaces: its
        // longest task against the run's and launch j's, with a launch's
        // fixed cost taken as launch_chunks chunks
        if (!ok || extra > max_extra || std::max(run_len, new_len) > run_len + own_len + launch_chunks) {
            run0 = j;
            run_len = 0;
            for (int t : tasks_of[j]) run_len = std::max(run_len, chunks(segs_of[t]));
            run_io.clear();
            note_run(j);
            continue;
        }
        joins[j] = 1;
        run_len = std::max(run_len, new_len);
        for (auto &tr : rewritten) {
            const int t = tr.first, m = L.tasks[t].m;
            std::vector<Seg> ns;
            for (const Piece &pc : tr.second) {
                if (pc.P < 0) {
                    ns.push_back(pc.sg);
                    continue;
                }
                const Seg &inner = segs_of[pc.P][pc.q];
                Seg s2 = inner;   // the producer segment's B operand, over the composite
                s2.a_buf = MAT_COMP;
                s2.a_off = (int64_t)L.comps.size();
                s2.a_trans = 0;
                s2.lda = m;
                L.comps.push_back(composite_of(pc.sg, inner, m));
                ns.push_back(s2);
            }
            segs_of[t].swap(ns);
        }
        note_run(j);
    }
    // segments in task order
    std::vector<Seg> segs;
    for (size_t t = 0; t < L.tasks.size(); ++t) {
        L.tasks[t].seg_begin = (int)segs.size();
        segs.insert(segs.end(), segs_of[t].begin(), segs_of[t].end());
        L.tasks[t].seg_end = (int)segs.size();
    }
    L.segs.swap(segs);
    // launches: a joining launch's tiles go to its run's launch
    std::vector<int2> tiles;
    std::vector<Launch> lns;
    std::vector<int> kinds;
    for (int li = 0; li < n; ++li) {
        if (!joins[li]) {
            lns.push_back(Launch{(int)tiles.size(), 0});
            kinds.push_back(L.kinds[li]);
        }
        const Launch &ln = L.launches[li];
        tiles.insert(tiles.end(), L.tiles.begin() + ln.tile_offset, L.tiles.begin() + ln.tile_offset + ln.ntiles);
        lns.back().ntiles = (int)tiles.size() - lns.back().tile_offset;
    }
    L.tiles.swap(tiles);
    L.launches.swap(lns);
    L.kinds.swap(kinds);
}

// A task with more than maxseg segments keeps its first maxseg; the rest go
// to continuation tasks (beta = 1, same output block and row tiles), the
// k-th continuation in a launch of its own inserted right after the task's
// launch, so the parts accumulate in order without write conflicts.  Used
// for the BF16x3 kernel, which stages a task's segment list in LDS.
void split_long_tasks(TaskLists &L, int maxseg) {
    std::vector<int2> tiles;
    std::vector<Launch> launches;
    std::vector<int> kinds;
    std::vector<int> seen(L.tasks.size(), -1);
    for (size_t li = 0; li < L.launches.size(); ++li) {
        const Launch ln = L.launches[li];
        // continuation tasks of this launch, by depth
        std::vector<std::vector<int>> cont;
        for (int t = ln.tile_offset; t < ln.tile_offset + ln.ntiles; ++t) {
            const int task = L.tiles[t].x;
            if (seen[task] == (int)li) continue;
            seen[task] = (int)li;
            const Task base = L.tasks[task];
            const int nseg = base.seg_end - base.seg_begin;
            if (nseg <= maxseg) continue;
            L.tasks[task].seg_end = base.seg_begin + maxseg;
            for (int k = 1; k * maxseg < nseg; ++k) {
                Task c = base;
                c.seg_begin = base.seg_begin + k * maxseg;
                c.seg_end = std::min(base.seg_end, c.seg_begin + maxseg);
                c.beta = 1;
                if ((int)cont.size() < k) cont.resize(k);
                cont[k - 1].push_back((int)L.tasks.size());
                L.tasks.push_back(c);
                L.task_kind.push_back(L.task_kind[task]);
            }
        }
        Launch nl{(int)tiles.size(), ln.ntiles};
        for (int t = ln.tile_offset; t < ln.tile_offset + ln.ntiles; ++t) tiles.push_back(L.tiles[t]);
        launches.push_back(nl);
        kinds.push_back(L.kinds[li]);
        for (auto &group : cont) {
            Launch cl{(int)tiles.size(), 0};
            for (int ct : group) {
                const int bm = L.task_kind[ct] == 1 ? X3_BM : SG_BM;
                for (int i = 0; i < (L.tasks[ct].m + bm - 1) / bm; ++i) tiles.push_back(make_int2(ct, i));
            }
            cl.ntiles = (int)tiles.size() - cl.tile_offset;
            launches.push_back(cl);
            kinds.push_back(L.kinds[li]);
        }
        seen.resize(L.tasks.size(), -1);
    }
    L.tiles.swap(tiles);
    L.launches.swap(launches);
    L.kinds.swap(kinds);
}

// Build the forward task lists for one kernel kind (0: segmented MFMA GEMM
// in FP64/FP32 with 64-row tiles, 1: split-BF16x3 with 128-row tiles).
// beta_u / beta_y select the accumulation rule of the coupling and leaf
// tasks (0 for mx_eval, the reference's rules for the drop-in entry points).
// wrow_in / skel_in give the row offsets of W blocks and skeleton blocks in
// the buffers the kernel reads (reference layout for kind 0, split panels for
// kind 1); Y always uses the reference layout.
TaskLists build_lists(const mx_structure &s, const Topology &t, int beta_u, int beta_y, int kind, int dir,
                      int sched, const int64_t *wrow_in, const int64_t *skel_in, bool group) {
    TaskLists L;
    const int nn = s.numnodes;
    // far pair k couples row node fx(k) with column node fy(k) through B_k
    auto fx = [&](int64_t k) { return s.is_hss ? (int)k + 1 : s.fpairx[k]; };
    auto fy = [&](int64_t k) {
        if (!s.is_hss) return s.fpairy[k];
        int x = (int)k + 1, p = t.parent[x];
        return s.lchildren[p] == x ? s.rchildren[p] : s.lchildren[p];
    };
    auto nx = [&](int64_t k) { return s.is_hss ? (int)k : s.npairx[k]; };
    auto ny = [&](int64_t k) { return s.is_hss ? (int)k : s.npairy[k]; };
    // pairs grouped by destination: the row node (forward) or the column
    // node (adjoint, where the block acts transposed)
    std::vector<std::vector<int64_t>> far_of(nn), near_of(s.nleaf);
    for (int64_t k = 0; k < s.fcount; ++k) far_of[dir == DIR_FWD ? fx(k) : fy(k)].push_back(k);
    for (int64_t k = 0; k < s.ncount; ++k) near_of[dir == DIR_FWD ? nx(k) : ny(k)].push_back(k);

    auto begin_launch = [&]() {
        L.launches.push_back(Launch{(int)L.tiles.size(), 0});
        L.kinds.push_back(kind);
    };
    auto end_launch = [&]() {
        L.launches.back().ntiles = (int)L.tiles.size() - L.launches.back().tile_offset;
        if (L.launches.back().ntiles == 0) {
            L.launches.pop_back();
            L.kinds.pop_back();
        }
    };
    auto new_task = [&](int64_t c_row, int m, int c_buf, int beta) {
        Task tk{};
        tk.c_row = c_row;
        tk.m = m;
        tk.c_buf = c_buf;
        tk.seg_begin = (int)L.segs.size();
        tk.seg_end = tk.seg_begin;
        tk.beta = beta;
        L.tasks.push_back(tk);
        L.task_kind.push_back(kind);
        return (int)L.tasks.size() - 1;
    };
    auto add_seg = [&](int task, int64_t a_off, int a_buf, int a_trans, int lda, int k, int b_buf, int64_t b_row,
                       int b_ld) {
        if (k <= 0) return;
        Seg sg{};
        sg.a_off = a_off;
        sg.a_buf = a_buf;
        sg.a_trans = a_trans;
        sg.lda = lda;
        sg.k = k;
        sg.b_buf = b_buf;
        sg.b_row = b_row;
        sg.b_ld = b_ld;
        L.segs.push_back(sg);
        L.tasks[task].seg_end = (int)L.segs.size();
    };
    const int maxd = (int)t.levels.size() - 1;
    // near pair k as a segment of the task of destination leaf lf
    auto near_src = [&](int64_t k) { return dir == DIR_FWD ? ny(k) : nx(k); };
    auto add_near = [&](int tk, int lf, int64_t k) {
        if (dir == DIR_FWD) {
            int y = ny(k);
            add_seg(tk, (int64_t)s.Dptr[k], MAT_D, 0, s.Ddim[lf], s.Ddim[y], VEC_W, wrow_in[y], s.Ddim[y]);
        } else {   // D_k^T, D_k stored Dim_x x Dim_y
            int x = nx(k);
            add_seg(tk, (int64_t)s.Dptr[k], MAT_D, 1, s.Ddim[x], s.Ddim[x], VEC_W, wrow_in[x], s.Ddim[x]);
        }
    };

    // Balance of the NEAR+LEAF launch (schedules with bit 1 set; the
    // autotuner decides whether it pays).  Its length is that of its longest
    // task; a leaf whose near list is much longer than the mean (the adjoint
    // groups near blocks by column, and one leaf can be near to many others)
    // would set it alone.  The first near segments of such a leaf move, as
    // partial tasks writing Y, into the UP launches of the upper levels,
    // which hold few tasks of bounded length: a piece is no longer than the
    // longest UP task of its launch and a launch takes at most 64 extra row
    // tiles, so neither grows much.  The first write of a leaf's Y keeps
    // beta_y; the later pieces and the NEAR+LEAF task accumulate (beta = 1).
    const int bk = kind == 1 ? X3_BK : SG_BK;
    const int bm = kind == 1 ? X3_BM : SG_BM;
    auto chunks = [&](int k) { return (k + bk - 1) / bk; };
    std::vector<size_t> nhoisted(s.nleaf, 0);                                  // leading near segments moved
    std::vector<std::vector<std::pair<int, std::pair<size_t, size_t>>>> pieces(maxd + 1);   // (leaf, [i0, i1))
    if (maxd >= 2 && (sched & 2)) {
        std::vector<int64_t> cost(s.nleaf, 0);
        int64_t total = 0, active = 0;
        for (int lf = 0; lf < s.nleaf; ++lf) {
            if (s.Ddim[lf] == 0) continue;
            for (int64_t k : near_of[lf]) cost[lf] += chunks(s.Ddim[near_src(k)]);
            const int v = t.node_of_leaf[lf];
            if (v > 0) cost[lf] += chunks(s.slen[v]);
            total += cost[lf];
            ++active;
        }
        const int64_t cap = active ? (5 * total + 4 * active - 1) / (4 * active) : 0;   // 1.25 x the mean
        std::vector<int> room(maxd + 1, 0), budget(maxd + 1, 0);
        for (int d = maxd - 1; d >= 1; --d) {
            int rows = 0;
            for (int v : t.levels[d]) {
                if (s.slen[v] == 0) continue;
                const int c = s.lchildren[v] == -1
                                  ? chunks(s.Ddim[s.lm[v]])
                                  : chunks(s.slen[s.lchildren[v]]) + chunks(s.slen[s.rchildren[v]]);
                room[d] = std::max(room[d], c);
                rows += (s.slen[v] + bm - 1) / bm;
            }
            budget[d] = std::max(0, 64 - rows);
        }
        std::vector<int> heavy;
        for (int lf = 0; lf < s.nleaf; ++lf)
            if (s.Ddim[lf] > 0 && cost[lf] > cap) heavy.push_back(lf);
        std::stable_sort(heavy.begin(), heavy.end(), [&](int a, int b) { return cost[a] > cost[b]; });
        for (int lf : heavy) {
            const int rt = (s.Ddim[lf] + bm - 1) / bm;
            size_t next = 0;
            for (int d = maxd - 1; d >= 1 && cost[lf] > cap; --d) {   // launch order
                if (budget[d] < rt) continue;
                size_t end = next;
                int c = 0;
                while (end < near_of[lf].size() && cost[lf] > cap) {
                    const int sc = chunks(s.Ddim[near_src(near_of[lf][end])]);
                    if (c + sc > room[d]) break;
                    c += sc;
                    cost[lf] -= sc;
                    ++end;
                }
                if (end == next) continue;
                pieces[d].push_back({lf, {next, end}});
                next = end;
                budget[d] -= rt;
            }
            nhoisted[lf] = next;
        }
    }
    std::vector<char> y_written(s.nleaf, 0);

    // UP: deepest level first
    for (int d = maxd; d >= 1; --d) {
        begin_launch();
        for (int v : t.levels[d]) {
            int sv = s.slen[v];
            if (sv == 0) continue;
            int tk = new_task(skel_in[v], sv, VEC_WSKEL, 0);
            if (s.lchildren[v] == -1) {
                int lf = s.lm[v];
                add_seg(tk, s.VTptr[v], MAT_VT, 0, sv, s.Ddim[lf], VEC_W, wrow_in[lf], s.Ddim[lf]);
            } else {
                int l = s.lchildren[v], r = s.rchildren[v];
                add_seg(tk, s.VTptr[v], MAT_VT, 0, sv, s.slen[l], VEC_WSKEL, skel_in[l], s.slen[l]);
                add_seg(tk, (int64_t)s.VTptr[v] + (int64_t)sv * s.slen[l], MAT_VT, 0, sv, s.slen[r], VEC_WSKEL,
                        skel_in[r], s.slen[r]);
            }
            add_tiles(L, tk, sv);
        }
        for (const auto &pc : pieces[d]) {   // near segments moved out of NEAR+LEAF
            const int lf = pc.first;
            int tk = new_task(s.wrow[lf], s.Ddim[lf], VEC_Y, y_written[lf] ? 1 : beta_y);
            for (size_t i = pc.second.first; i < pc.second.second; ++i) add_near(tk, lf, near_of[lf][i]);
            y_written[lf] = 1;
            add_tiles(L, tk, s.Ddim[lf]);
        }
        end_launch();
    }
    auto add_far = [&](int tk, int c) {
        const int sc = s.slen[c];
        for (int64_t k : far_of[c]) {
            if (dir == DIR_FWD) {
                int y = fy(k);
                add_seg(tk, (int64_t)s.Bptr[k], MAT_B, 0, sc, s.slen[y], VEC_WSKEL, skel_in[y], s.slen[y]);
            } else {   // B_k^T, B_k stored s_x x s_y
                int x = fx(k);
                add_seg(tk, (int64_t)s.Bptr[k], MAT_B, 1, s.slen[x], s.slen[x], VEC_WSKEL, skel_in[x], s.slen[x]);
            }
        }
    };
    auto add_down = [&](int tk, int c) {   // VT_p[:, part(c)]^T u_p; false when there is none
        int p = t.parent[c];
        if (p <= 0 || s.slen[p] == 0) return false;
        int sp = s.slen[p];
        int64_t off = s.VTptr[p];
        if (s.rchildren[p] == c) off += (int64_t)sp * s.slen[s.lchildren[p]];
        add_seg(tk, off, MAT_VT, 1, sp, sp, VEC_USKEL, skel_in[p], sp);
        return true;
    };
    if ((sched & 1) == 0) {
        // FAR + DOWN, root's children first: u_c = sum_far B w_y + VT_p[:, part(c)]^T u_p
        for (int d = 1; d <= maxd; ++d) {
            begin_launch();
            for (int c : t.levels[d]) {
                int sc = s.slen[c];
                if (sc == 0) continue;
                int tk = new_task(skel_in[c], sc, VEC_USKEL, beta_u);
                add_far(tk, c);
                add_down(tk, c);
                add_tiles(L, tk, sc);
            }
            end_launch();
        }
    } else {
        // FAR for every level at once, then DOWN level by level (beta = 1)
        begin_launch();
        for (int d = 1; d <= maxd; ++d)
            for (int c : t.levels[d]) {
                int sc = s.slen[c];
                if (sc == 0) continue;
                int tk = new_task(skel_in[c], sc, VEC_USKEL, beta_u);
                add_far(tk, c);
                add_tiles(L, tk, sc);
            }
        end_launch();
        for (int d = 2; d <= maxd; ++d) {
            begin_launch();
            for (int c : t.levels[d]) {
                int sc = s.slen[c];
                if (sc == 0) continue;
                int tk = new_task(skel_in[c], sc, VEC_USKEL, 1);
                if (add_down(tk, c)) add_tiles(L, tk, sc);
                else L.tasks.pop_back();
            }
            end_launch();
        }
    }
    // NEAR + LEAF: Y_x = sum_near D W_y + VT_x^T u_x (less the near segments
    // moved into the UP launches)
    begin_launch();
    for (int lf = 0; lf < s.nleaf; ++lf) {
        if (s.Ddim[lf] == 0) continue;
        int tk = new_task(s.wrow[lf], s.Ddim[lf], VEC_Y, y_written[lf] ? 1 : beta_y);
        for (size_t i = nhoisted[lf]; i < near_of[lf].size(); ++i) add_near(tk, lf, near_of[lf][i]);
        int v = t.node_of_leaf[lf];
        int sv = s.slen[v];
        if (v > 0 && sv > 0) add_seg(tk, s.VTptr[v], MAT_VT, 1, sv, sv, VEC_USKEL, skel_in[v], sv);
        add_tiles(L, tk, s.Ddim[lf]);
    }
    end_launch();
    // at most 4.5 MFLOP per right-hand side column for each merge
    if (sched_coarse(sched)) coarsen(L, 4.5e6, X3_MAXSEG, X3_BK, 4);
    order_tiles(L, kind == 1 ? X3_BK : SG_BK, kind == 1 || group);
    return L;
}

int64_t sweep_ctl_ints(int numnodes, int nrhs) {
    return SWEEP_CNT + (int64_t)2 * numnodes * ((nrhs + 31) / 32);
}

SweepLists build_sweep(const TaskLists &X, int numnodes, const std::vector<int64_t> &sprow) {
    SweepLists S;
    S.nblocks = 2 * numnodes;
    std::map<int64_t, int> node_at;   // first panel row -> skeleton node
    for (int v = 0; v < numnodes; ++v) node_at[sprow[v]] = v;
    // counter block of a panel block: w^ of node v is v, u^ of node v is numnodes + v
    auto block_of = [&](int buf, int64_t row) {
        auto it = node_at.find(row);
        if (it == node_at.end()) throw MxError(MX_ERR_ARG, "internal: a skeleton block without a node");
        return buf == VEC_WSKEL ? it->second : numnodes + it->second;
    };
    const int ntask = (int)X.tasks.size();
    auto rowtiles = [&](int t) { return (X.tasks[t].m + X3_BM - 1) / X3_BM; };
    // the tasks of the launches, and the one writer of every output block
    std::vector<char> used(ntask, 0);
    for (const int2 &tl : X.tiles)
        if (tl.x != X.dummy) used[tl.x] = 1;
    std::vector<int> writer(S.nblocks, -1);
    std::set<int64_t> y_rows;
    for (int t = 0; t < ntask; ++t) {
        if (!used[t]) continue;
        const Task &tk = X.tasks[t];
        if (tk.c_buf == VEC_Y) {
            if (!y_rows.insert(tk.c_row).second) return S;   // several writers of a Y block
        } else {
            const int b = block_of(tk.c_buf, tk.c_row);
            if (writer[b] >= 0) return S;                      // several writers of a skeleton block
            writer[b] = t;
        }
    }
    // tasks with their u^ segments last, and their waits
    S.tasks = X.tasks;
    S.stasks.assign(ntask, SweepTask{0, 0, 0, 0, -1, -1, {0, 0}});
    for (int t = 0; t < ntask; ++t) {
        Task &tk = S.tasks[t];
        const int b0 = (int)S.segs.size();
        std::vector<Seg> early, late;
        for (int i = X.tasks[t].seg_begin; i < X.tasks[t].seg_end && t != X.dummy; ++i)
            (X.segs[i].b_buf == VEC_USKEL ? late : early).push_back(X.segs[i]);
        SweepTask &st = S.stasks[t];
        auto add_waits = [&](const std::vector<Seg> &segs) {
            std::set<int> seen;
            for (const Seg &g : segs) {
                if (g.b_buf != VEC_WSKEL && g.b_buf != VEC_USKEL) continue;
                const int b = block_of(g.b_buf, g.b_row);
                if (!seen.insert(b).second) continue;
                if (writer[b] < 0) throw MxError(MX_ERR_ARG, "internal: a skeleton block read but never written");
                S.waits.push_back(SweepWait{b, SWEEP_WAVES * rowtiles(writer[b])});
            }
        };
        st.ew_b = (int)S.waits.size();
        add_waits(early);
        st.ew_e = (int)S.waits.size();
        st.lw_b = st.ew_e;
        add_waits(late);
        st.lw_e = (int)S.waits.size();
        st.late_seg = late.empty() ? -1 : (int)early.size();
        st.sig = tk.c_buf == VEC_Y || t == X.dummy ? -1 : block_of(tk.c_buf, tk.c_row);
        for (const Seg &g : early) S.segs.push_back(g);
        for (const Seg &g : late) S.segs.push_back(g);
        tk.seg_begin = b0;
        tk.seg_end = (int)S.segs.size();
    }
    // queues: the XCD groups of every launch (eight contiguous parts, as
    // xcd_remap deals them), launch by launch; rank g interleaves the
    // groups position by position
    std::vector<std::vector<SweepTile>> q(SWEEP_NQ);
    std::map<std::pair<int, int>, std::pair<int, int>> at;   // (task, row tile) -> (queue, index)
    int gbase = 0;
    for (const Launch &ln : X.launches) {
        const int n = ln.ntiles, per = n / SWEEP_NQ, rem = n % SWEEP_NQ;
        int longest = 0;
        for (int z = 0; z < SWEEP_NQ; ++z) {
            const int b = z * per + std::min(z, rem), e = b + per + (z < rem ? 1 : 0);
            int p = 0;
            for (int i = b; i < e; ++i) {
                const int2 tl = X.tiles[ln.tile_offset + i];
                if (tl.x == X.dummy) continue;   // (a task without segments still writes zeros)
                SweepTile st{};
                st.task = tl.x;
                st.rt = tl.y;
                st.g = gbase + p * SWEEP_NQ + z;
                at[{tl.x, tl.y}] = {z, (int)q[z].size()};
                q[z].push_back(st);
                ++p;
            }
            longest = std::max(longest, p);
        }
        gbase += longest * SWEEP_NQ;
    }
    // producers: every row tile of the writer of each block a task reads;
    // each must rank below the consumer (a producer is in an earlier launch)
    for (int z = 0; z < SWEEP_NQ; ++z)
        for (SweepTile &st : q[z]) {
            for (int y = 0; y < SWEEP_NQ; ++y) st.maxtile[y] = -1;
            const SweepTask &sk = S.stasks[st.task];
            for (int wi = sk.ew_b; wi < sk.lw_e; ++wi) {
                const int p = writer[S.waits[wi].block];
                for (int r = 0; r < rowtiles(p); ++r) {
                    auto it = at.find({p, r});
                    if (it == at.end()) throw MxError(MX_ERR_ARG, "internal: a producer tile outside the launches");
                    const int y = it->second.first, j = it->second.second;
                    if (q[y][j].g >= st.g) throw MxError(MX_ERR_ARG, "internal: a producer ranked after its consumer");
                    st.maxtile[y] = std::max(st.maxtile[y], j);
                }
            }
        }
    for (int z = 0; z < SWEEP_NQ; ++z) {
        S.qinfo.push_back(make_int2((int)S.tiles.size(), (int)q[z].size()));
        S.tiles.insert(S.tiles.end(), q[z].begin(), q[z].end());
    }
    S.ok = true;
    return S;
}

// Row tile r of a task maps to row pair r / 2; each pair appears once per
// launch, where its first tile was.  A launch that order_tiles cut into
// eight padded XCD groups keeps that cut (each group re-padded to the
// longest).
void pair_tiles(TaskLists &L) {
    L.tiles2.clear();
    L.launches2.clear();
    for (const Launch &ln : L.launches) {
        const int ng = (ln.ntiles % 8 == 0) ? 8 : 1;
        const int per = ln.ntiles / ng;
        std::vector<std::vector<int2>> grp(ng);
        std::set<std::pair<int, int>> seen;
        for (int q = 0; q < ng; ++q) {
            for (int i = 0; i < per; ++i) {
                const int2 t = L.tiles[ln.tile_offset + q * per + i];
                if (t.x == L.dummy) continue;
                if (seen.insert({t.x, t.y / 2}).second) grp[q].push_back(make_int2(t.x, t.y / 2));
            }
        }
        size_t gmax = 0;
        for (auto &g : grp) gmax = std::max(gmax, g.size());
        Launch nl{(int)L.tiles2.size(), (int)(gmax * ng)};
        for (auto &g : grp) {
            for (const int2 &p : g) L.tiles2.push_back(p);
            for (size_t i = g.size(); i < gmax; ++i) L.tiles2.push_back(make_int2(L.dummy, 0));
        }
        L.launches2.push_back(nl);
    }
}

// Gradient products of every stored block (see mx_grad.h): one task per
// near block, far block, leaf VT and half of an internal VT, tiled 64 x 64.

GradLists build_grad_lists(const mx_structure &s, const Topology &t) {
    GradLists G;
    auto task = [&](int64_t c_off, int m, int n, int ldc, int c_buf) {
        GTask k{};
        k.c_off = c_off;
        k.m = m;
        k.n = n;
        k.ldc = ldc;
        k.c_buf = c_buf;
        k.seg_begin = k.seg_end = (int)G.segs.size();
        G.tasks.push_back(k);
        return (int)G.tasks.size() - 1;
    };
    auto seg = [&](int tk, int x_buf, int64_t x_row, int x_ld, int z_buf, int64_t z_row, int z_ld) {
        GSeg g{};
        g.x_buf = x_buf;
        g.x_row = x_row;
        g.x_ld = x_ld;
        g.z_buf = z_buf;
        g.z_row = z_row;
        g.z_ld = z_ld;
        G.segs.push_back(g);
        G.tasks[tk].seg_end = (int)G.segs.size();
        G.flops += 2.0 * x_ld * z_ld;
    };
    auto tiles = [&](int tk) {
        const GTask &k = G.tasks[tk];
        for (int i = 0; i < (k.m + 63) / 64; ++i)
            for (int j = 0; j < (k.n + 63) / 64; ++j) G.tiles.push_back(make_int4(tk, i, j, 0));
        constexpr int T = GRAD_X3_TILE, TW = GRAD_X3_WTILE;
        for (int i = 0; i < (k.m + T - 1) / T; ++i)
            for (int j = 0; j < (k.n + T - 1) / T; ++j) G.tiles_x3.push_back(make_int4(tk, i, j, 0));
        for (int i = 0; i < (k.m + TW - 1) / TW; ++i)
            for (int j = 0; j < (k.n + TW - 1) / TW; ++j) G.tiles_x3w.push_back(make_int4(tk, i, j, 0));
        G.area_x3 += (double)((k.m + T - 1) / T) * ((k.n + T - 1) / T) * T * T * (k.seg_end - k.seg_begin);
        G.area_x3w += (double)((k.m + TW - 1) / TW) * ((k.n + TW - 1) / TW) * TW * TW * (k.seg_end - k.seg_begin);
    };
    for (int64_t k = 0; k < s.ncount; ++k) {
        int x = s.is_hss ? (int)k : s.npairx[k], y = s.is_hss ? (int)k : s.npairy[k];
        int mx_ = s.Ddim[x], ny_ = s.Ddim[y];
        if (mx_ == 0 || ny_ == 0) continue;
        int tk = task((int64_t)s.Dptr[k], mx_, ny_, mx_, MAT_D);
        seg(tk, GB_DY, s.wrow[x], mx_, GB_W, s.wrow[y], ny_);
        tiles(tk);
    }
    for (int64_t k = 0; k < s.fcount; ++k) {
        int x, y;
        if (s.is_hss) {
            x = (int)k + 1;
            int p = t.parent[x];
            y = s.lchildren[p] == x ? s.rchildren[p] : s.lchildren[p];
        } else {
            x = s.fpairx[k];
            y = s.fpairy[k];
        }
        int sx = s.slen[x], sy = s.slen[y];
        if (sx == 0 || sy == 0) continue;
        int tk = task((int64_t)s.Bptr[k], sx, sy, sx, MAT_B);
        seg(tk, GB_YHAT, s.skelrow[x], sx, GB_WHAT, s.skelrow[y], sy);
        tiles(tk);
    }
    for (int v = 1; v < s.numnodes; ++v) {
        int sv = s.slen[v];
        if (sv == 0) continue;
        if (s.lchildren[v] == -1) {
            int lf = s.lm[v], dm = s.Ddim[lf];
            if (dm == 0) continue;
            int tk = task(s.VTptr[v], sv, dm, sv, MAT_VT);
            seg(tk, GB_GHAT, s.skelrow[v], sv, GB_W, s.wrow[lf], dm);
            seg(tk, GB_UHAT, s.skelrow[v], sv, GB_DY, s.wrow[lf], dm);
            tiles(tk);
        } else {
            int64_t off = s.VTptr[v];
            for (int c : {s.lchildren[v], s.rchildren[v]}) {
                int sc = s.slen[c];
                if (sc > 0) {
                    int tk = task(off, sv, sc, sv, MAT_VT);
                    seg(tk, GB_GHAT, s.skelrow[v], sv, GB_WHAT, s.skelrow[c], sc);
                    seg(tk, GB_UHAT, s.skelrow[v], sv, GB_YHAT, s.skelrow[c], sc);
                    tiles(tk);
                }
                off += (int64_t)sv * sc;
            }
        }
    }
    // xcd_remap gives XCD x the x-th contiguous eighth of the tiles.  Each
    // gradient type (D, B, VT: their tiles sit in that order, VT tiles with
    // twice the segments) is cut into eight contiguous parts of equal work,
    // and eighth x holds part x of every type, longest tiles first, padded
    // with empty tiles to the longest eighth: the work is even across the
    // XCDs whichever gradients a call requests, and each XCD still walks
    // neighbouring blocks.
    GTask empty{};
    empty.ldc = 1;
    empty.c_buf = MAT_D;
    empty.seg_begin = empty.seg_end = 0;
    const int dummy = (int)G.tasks.size();
    G.tasks.push_back(empty);
    auto balance = [&](std::vector<int4> &tiles) {
        const int ng = 8;
        std::vector<std::vector<int4>> eighth(ng);
        for (int type : {MAT_VT, MAT_D, MAT_B}) {
            std::vector<int4> v;
            int64_t total = 0;
            for (const int4 &tl : tiles)
                if (G.tasks[tl.x].c_buf == type) {
                    v.push_back(tl);
                    total += G.tasks[tl.x].seg_end - G.tasks[tl.x].seg_begin;
                }
            int64_t acc = 0;
            int q = 0;
            for (const int4 &tl : v) {
                const int64_t c = G.tasks[tl.x].seg_end - G.tasks[tl.x].seg_begin;
                while (q < ng - 1 && (acc + c / 2) * ng >= total * (q + 1)) ++q;   // part by the tile's midpoint
                eighth[q].push_back(tl);
                acc += c;
            }
        }
        size_t longest = 0;
        for (const auto &e : eighth) longest = std::max(longest, e.size());
        tiles.clear();
        for (const auto &e : eighth) {
            tiles.insert(tiles.end(), e.begin(), e.end());
            for (size_t i = e.size(); i < longest; ++i) tiles.push_back(make_int4(dummy, 0, 0, 0));
        }
    };
    balance(G.tiles);
    balance(G.tiles_x3);
    balance(G.tiles_x3w);
    return G;
}

}  // namespace mx
