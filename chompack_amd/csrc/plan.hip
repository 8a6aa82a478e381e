// Device plan: level schedule + work lists for every sweep, built once per
// symbolic factorization, and the C-ABI entry points (include/chompack_amd.h).
//
// A plan copies the symbolic arrays into HBM and precomputes, per level of
// the supernodal elimination tree (level = height above the leaves), the
// launch sequence of each sweep:
//   * small fronts (LDS-resident)  -> one k_*_small launch per level;
//   * large fronts                 -> column-tiled extend-add, 64-wide panel
//     steps (diagonal factor + panel solve + MFMA trailing update), and the
//     stored-convention solve, each step one batched launch for all large
//     fronts of the level.
// Executing a sweep replays the precomputed launches on the caller's stream;
// there is no host work per front, only the final status read.
#include "plan.h"
#include <hip/hip_ext.h>
#include "../../include/chompack_amd.h"

#include <algorithm>
#include <map>
#include <set>
#include <cmath>
#include <array>
#include <cstring>
#include <functional>
#include <queue>

static thread_local std::string g_err;
void chp_set_error(const std::string& msg) { g_err = msg; }

namespace chpi {

size_t chol_small_lds(const FrontInfo& f) {
    size_t nj = f.nn + f.na;
    return nj * nj * sizeof(double);
}
size_t projinv_small_lds(const FrontInfo& f) {
    size_t nn = f.nn, na = f.na;
    return (2 * nn * nn + 2 * na * nn + nn + na * na) * sizeof(double);
}
size_t llt_small_lds(const FrontInfo& f) {
    size_t nn = f.nn, nj = f.nn + f.na;
    return (nj * nn + nj * nj) * sizeof(double);
}
bool chol_is_small(const chp_plan* P, const FrontInfo& f) {
    return f.nn + f.na <= P->small_max && chol_small_lds(f) <= P->lds_cap;
}
bool projinv_is_small(const chp_plan* P, const FrontInfo& f) {
    return f.nn + f.na <= P->small_max && projinv_small_lds(f) <= P->lds_cap;
}

GemmTask mk(int bufA, i64 offA, int lda, int opA, int bufB, i64 offB, int ldb, int opB,
            int bufC, i64 offC, int ldc, int M, int N, int K, double alpha, double beta,
            int tri, int triA, int triB) {
    GemmTask t;
    std::memset(&t, 0, sizeof(t));
    t.offA = offA; t.offB = offB; t.offC = offC;
    t.lda = lda; t.ldb = ldb; t.ldc = ldc;
    t.M = M; t.N = N; t.K = K;
    t.bufA = (unsigned char)bufA; t.bufB = (unsigned char)bufB; t.bufC = (unsigned char)bufC;
    t.opA = (unsigned char)opA; t.opB = (unsigned char)opB;
    t.tri = (unsigned char)tri; t.triA = (unsigned char)triA; t.pad = (unsigned char)triB;
    t.alpha = alpha; t.beta = beta;
    t.ksplit = 1; t.kchunk = K; t.offP = 0;
    return t;
}

// Useful flops of a task: 2 * sum over computed C(i,j) of the number of k
// with nonzero opA(i,k) opB(k,j) (triangular operands and outputs excluded).
double task_flops(const GemmTask& t) {
    double total = 0.0;
    for (int i = 0; i < t.M; ++i) {
        int jmax = t.tri ? std::min(i + 1, t.N) : t.N;
        if (jmax <= 0) continue;
        int klo = 0, khi = t.K;
        if (t.triA == 1) khi = std::min(khi, i + 1);
        if (t.triA == 2) klo = std::max(klo, i);
        if (khi <= klo) continue;
        if (t.pad == 1) {  // B lower triangular: k >= j
            double sum = (double)std::min(jmax, klo) * (khi - klo);
            int a = klo, b = std::min(jmax, khi);
            if (b > a) sum += (double)(b - a) * khi - ((double)(b - 1) * b / 2.0 - (double)(a - 1) * a / 2.0);
            total += sum;
        } else {
            total += (double)jmax * (khi - klo);
        }
    }
    return 2.0 * total;
}

// Adds the tiles of one task.  Tasks with few output tiles and a long K are
// split along K (deterministic split-K) so the launch fills the GPU.
// Tasks k_gemm_narrow takes: at most NARROW_N columns, op(A) plain or
// triangular, op(B) plain (op(B) = B with a triangular A).
bool narrow_task(const GemmTask& t) {
    return t.N <= NARROW_N && t.opA <= 1 && t.opB <= 1 && t.tri == 0 && t.pad == 0 &&
           (t.triA == 0 || t.opB == 0);
}

void add_gemm_tiles(chp_plan* P, std::vector<GemmTile>& tiles, GemmTask t, bool skip_upper) {
    if (t.M <= 0 || t.N <= 0) return;
    int id = (int)P->pg->tasks.size();
    int mt = (t.M + TILE - 1) / TILE, nt = (t.N + TILE - 1) / TILE;
    int ntiles = 0;
    for (int tn = 0; tn < nt; ++tn)
        for (int tm = 0; tm < mt; ++tm)
            if (!(skip_upper && tm < tn)) ++ntiles;
    int ks = 1;
    if (P->splitk_max <= 1) {
        ks = 1;
    } else if (narrow_task(t) && !skip_upper) {
        // narrow tiles are cheap: slices of about 256 k (64 per wave)
        ks = std::max(1, std::min(t.K / 256, std::min(16, P->splitk_max)));
    } else if (ntiles < 128 && t.K >= 512) {
        ks = std::max(1, std::min(t.K / 256, (512 + ntiles - 1) / ntiles));
        ks = std::min(ks, std::min(16, P->splitk_max));
    } else if (ntiles < 1024 && t.K >= 768) {
        // a few tiles per CU: split up to 4 ways so that the launch holds
        // about 2048 units of K >= 384 (the tail of a 2-3 tile-per-CU launch
        // otherwise leaves CUs idle)
        ks = std::max(1, std::min(t.K / 384, (2048 + ntiles - 1) / ntiles));
        ks = std::min(ks, std::min(4, P->splitk_max));
    }
    if (ks > 1) {
        int chunk = (t.K + ks - 1) / ks;
        chunk = (chunk + 31) / 32 * 32;
        ks = (t.K + chunk - 1) / chunk;
        t.ksplit = ks;
        t.kchunk = chunk;
        Program& pg = *P->pg;
        if (pg.split_side) {
            t.offP = pg.side_cur;
            pg.side_cur += (i64)ks * t.M * t.N;
            pg.side_len = std::max(pg.side_len, pg.side_cur);
            pg.side_tasks.push_back(id);
        } else {
            t.offP = pg.splitp_cur;
            pg.splitp_cur += (i64)ks * t.M * t.N;
            pg.splitp_len = std::max(pg.splitp_len, pg.splitp_cur);
        }
    }
    P->pg->tasks.push_back(t);
    for (int tn = 0; tn < nt; ++tn)
        for (int tm = 0; tm < mt; ++tm) {
            if (skip_upper && tm < tn) continue;
            for (int k = 0; k < ks; ++k) tiles.push_back({id, tm, tn, k});
            if (ks > 1) P->pg->split_out.push_back({id, tm, tn, 0});
        }
}

// True when C's storage overlaps A's or B's (in-place products such as the
// solves by an explicit inverse): such tasks need every operand read before
// any output is written, which the tile kernels do and k_gemm_small does not.
static bool c_aliases(const GemmTask& t) {
    auto span = [](i64 off, int ld, int rows, int cols) {
        return std::make_pair(off, off + (i64)ld * std::max(cols - 1, 0) + rows);
    };
    const auto c = span(t.offC, t.ldc, t.M, t.N);
    const auto a = t.opA == 1 ? span(t.offA, t.lda, t.K, t.M) : span(t.offA, t.lda, std::max(t.M, t.K), t.K);
    const auto b = t.opB == 1 ? span(t.offB, t.ldb, t.N, t.K) : span(t.offB, t.ldb, std::max(t.K, t.N), t.N);
    auto hit = [&](std::pair<i64, i64> o) { return c.first < o.second && o.first < c.second; };
    return (t.bufA == t.bufC && hit(a)) || (t.bufB == t.bufC && hit(b));
}

void push_gemm(chp_plan* P, std::vector<Launch>& seq, const std::vector<GemmTile>& all,
               int schur) {
    if (all.empty()) return;
    // tasks far smaller than a tile take one wave each (k_gemm_small) instead
    // of a padded 64 x 64 MFMA tile; a task of at most 64 x 64 has one tile
    std::vector<GemmTile> tiles;
    std::vector<int> small;
    double fls = 0.0;
    for (const GemmTile& g : all) {
        const GemmTask& t = P->pg->tasks[g.task];
        // (the low-precision sweep keeps its Schur updates on the low-precision tiles)
        if (t.ksplit == 1 && t.M <= TILE && t.N <= TILE && t.K <= SMALL_TASK_K && !c_aliases(t) &&
            !(schur && P->building_lowprec)) {
            small.push_back(g.task);
            fls += task_flops(t);
        } else {
            tiles.push_back(g);
        }
    }
    // a few tiny tasks beside regular tiles stay in the tile launch (a launch
    // of their own would add a step to the sweep's critical path)
    if (!small.empty() && !tiles.empty() && small.size() < 256) {
        tiles = all;
        small.clear();
    }
    if (!small.empty()) {
        Launch S{K_GEMM_SMALL, (int)small.size(), 256, 0, P->pg->lists.add(small),
                 (int)small.size()};
        S.flops = fls;
        seq.push_back(S);
    }
    if (tiles.empty()) return;
    P->n_gemm_tiles += (i64)tiles.size();
    // launches too small to fill the device take the latency-oriented tile
    // kernel (512 threads), split-K slices included
    bool narrow = !schur;
    for (const GemmTile& g : tiles) narrow = narrow && narrow_task(P->pg->tasks[g.task]);
    const int block = (!narrow && (int)tiles.size() <= LAT_TILES) ? 512 : 256;
    Launch L{narrow ? K_GEMM_NARROW : K_GEMM, (int)tiles.size(), block, 0, P->pg->lists.add(tiles),
             schur};
    double fl = 0.0;
    int last = -1;
    for (const GemmTile& g : tiles)
        if (g.task != last) { fl += task_flops(P->pg->tasks[g.task]); last = g.task; }
    L.flops = fl;
    seq.push_back(L);
    // the split tasks whose tiles are in this list get their reduction now
    std::vector<GemmTile> red;
    std::vector<GemmTile> keep;
    for (const GemmTile& g : P->pg->split_out) {
        bool mine = false;
        for (const GemmTile& h : tiles) if (h.task == g.task) { mine = true; break; }
        (mine ? red : keep).push_back(g);
    }
    P->pg->split_out.swap(keep);
    if (!red.empty())  // eight workgroups per output tile
        seq.push_back({K_SPLITK_REDUCE, 8 * (int)red.size(), 256, 0, P->pg->lists.add(red), 0});
    if (P->pg->split_out.empty()) P->pg->splitp_cur = P->pg->side_cur = 0;
}

std::vector<int4> extadd_tiles(const chp_plan* P, const std::vector<int>& fronts, int* ntiles) {
    std::vector<int4> et, bnd;
    for (int k : fronts) {
        const FrontInfo& f = P->fi[k];
        const int nj = f.nn + f.na;
        for (int c0 = 0; c0 < nj; c0 += 16)
            for (int r0 = c0 & ~255; r0 < nj; r0 += 256) {
                // the children's index ranges landing in the tile (first child
                // index with relative index >= c0, c1, r0, r1), children in
                // reverse order
                et.push_back(make_int4(k, c0, r0, (int)bnd.size()));
                const int v[4] = {c0, std::min(c0 + 16, nj), r0, std::min(r0 + 256, nj)};
                for (int l = f.chend - 1; l >= f.chbeg; --l) {
                    const FrontInfo& g = P->fi[P->h_chidx[l]];
                    const int* r = P->h_relidx.data() + g.relptr;
                    int b[4];
                    for (int q = 0; q < 4; ++q) b[q] = (int)(std::lower_bound(r, r + g.na, v[q]) - r);
                    bnd.push_back(make_int4(b[0], b[1], b[2], b[3]));
                    // and the child's own fields, so the kernel needs no
                    // dependent loads of chidx and FrontInfo
                    bnd.push_back(make_int4(g.na, 0, (int)(g.relptr & 0xffffffff), (int)(g.relptr >> 32)));
                    bnd.push_back(make_int4((int)(g.updptr & 0xffffffff), (int)(g.updptr >> 32), 0, 0));
                }
            }
    }
    *ntiles = (int)et.size();
    for (int4& t : et) t.w += *ntiles;  // the bounds follow the tiles in the list
    et.insert(et.end(), bnd.begin(), bnd.end());
    return et;
}

std::vector<int> children_of(const chp_plan* P, const std::vector<int>& fr) {
    std::vector<int> ch;
    for (int k : fr)
        for (int l = P->fi[k].chbeg; l < P->fi[k].chend; ++l) ch.push_back(P->h_chidx[l]);
    return ch;
}

std::vector<int4> extract_tiles(const chp_plan* P, const std::vector<int>& fronts) {
    std::vector<int4> xt;
    for (int k : fronts) {
        const int na = P->fi[k].na;
        for (int j0 = 0; j0 < na; j0 += 16)
            for (int i0 = j0 & ~255; i0 < na; i0 += 256) xt.push_back(make_int4(k, j0, i0, 0));
    }
    return xt;
}

// A persistent launch of a small-front kernel over `fronts`: the list in
// decreasing order of work (nj^2 (nn + 1), largest first, so the longest
// fronts start first and the tail is short), and at most as many
// workgroups as the device holds at once with `lds` bytes each; iarg
// carries the front count.
Launch small_launch(chp_plan* P, Kid kid, std::vector<int> fronts, size_t lds) {
    std::stable_sort(fronts.begin(), fronts.end(), [&](int a, int b) {
        const FrontInfo &fa = P->fi[a], &fb = P->fi[b];
        const double ja = fa.nn + fa.na, jb = fb.nn + fb.na;
        return ja * ja * (fa.nn + 1) > jb * jb * (fb.nn + 1);
    });
    const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(8, P->lds_per_cu / std::max<size_t>(lds, 1)));
    const int count = (int)fronts.size();
    const int grid = std::min(count, P->ncu * per_cu);
    return Launch{kid, grid, 256, lds, P->pg->lists.add(fronts), count};
}

// Chains: runs of at least CHAIN_MIN consecutive levels that hold one small
// front each (an elimination-tree path, as in banded and arrow patterns), of
// order at most CHAIN_NJ, go to one launch of a single workgroup that walks
// the fronts in order: child before parent for cholesky and llt, parent
// before child for the projected inverse.  Levels are heights above the
// leaves, so the one front of a level is a child of the next level's one
// front; the matrix handed from one to the other stays in LDS (keep / nbr of
// ChainFront).  A level schedule would spend launches per front there.
constexpr int CHAIN_MIN = 3;

Launch chain_launch(chp_plan* P, Kid kid, const std::vector<int>& fronts, bool child_first) {
    const size_t n = fronts.size();
    std::vector<ChainFront> cf(n);
    for (size_t q = 0; q < n; ++q) {
        const int k = fronts[q];
        const FrontInfo& f = P->fi[k];
        ChainFront& c = cf[q];
        std::memset(&c, 0, sizeof(c));
        c.blkptr = f.blkptr;
        c.updptr = f.updptr;
        c.front = k;
        c.nn = f.nn;
        c.na = f.na;
        c.chbeg = f.chbeg;
        c.chend = f.chend;
        c.nbr = -1;
        if (q + 1 < n)
            c.keep = child_first ? (f.parent == fronts[q + 1]) : (P->fi[fronts[q + 1]].parent == k);
        if (q > 0 && cf[q - 1].keep) c.nbr = fronts[q - 1];
        // the relative indices of the child handed over in LDS
        const int ch = child_first ? (c.nbr >= 0 ? c.nbr : -1) : (c.keep ? fronts[q + 1] : -1);
        if (ch >= 0) {
            c.nbr_na = P->fi[ch].na;
            c.nbr_relptr = P->fi[ch].relptr;
        }
    }
    return Launch{kid, 1, CHAIN_TPB, 0, P->pg->lists.add(cf), (int)n};
}

// Tile-DAG launch for the large fronts of one level (k_chol_dag).  The tasks
// of each front are generated step by step (POTRF(k), TRSM(., k), UPDATE(.,
// ., k)), each with its dependencies: the previous operation on its tile and
// the operations that produce its operand tiles.  The queue order is the
// start order of a list schedule on as many workers as the launch has
// workgroups, in which a free worker takes the ready task with the longest
// path to the end of the DAG (estimated durations): a topological order that
// starts the pivot chain's tasks as soon as they are ready.
namespace {
double dag_cost(int type, int m, int n, int k) {
    // microseconds, from the task shapes (fixed costs of the waits and fences
    // included); only the relative sizes matter for the order
    const double f = (double)m * n * k / (64.0 * 64.0 * 64.0);
    if (type == DAG_POTRF) return 10.0;
    return 1.5 + 1.5 * f;
}
}  // namespace

void build_chol_dag(chp_plan* P, const std::vector<int>& large, std::vector<Launch>& seq) {
    std::vector<DagFront> fr;
    std::vector<DagTask> tk;
    std::vector<double> dur, fl;
    std::vector<int> dptr{0}, dep;  // dependencies of task t: dep[dptr[t] .. dptr[t + 1])
    for (int key : large) {
        const FrontInfo& f = P->fi[key];
        DagFront F;
        std::memset(&F, 0, sizeof(F));
        F.blkptr = f.blkptr; F.updptr = f.updptr; F.dinvoff = P->h_dinvptr[key];
        F.nn = f.nn; F.na = f.na; F.key = key;
        F.P = (f.nn + NB - 1) / NB;
        F.R = F.P + (f.na + NB - 1) / NB;
        F.cbase = (int)P->dagcnt_len;
        P->dagcnt_len += (i64)F.R * (F.R + 1) / 2;
        const int fidx = (int)fr.size();
        fr.push_back(F);
        const int R = F.R;
        auto rows = [&](int i) {
            return i < F.P ? std::min(NB, f.nn - NB * i) : std::min(NB, f.na - NB * (i - F.P));
        };
        std::vector<int> last((size_t)R * (R + 1) / 2, -1);
        auto at = [&](int i, int j) -> int& { return last[(size_t)i * (i + 1) / 2 + j]; };
        auto add = [&](int type, int i, int j, int k, std::initializer_list<int> deps) {
            const int t = (int)tk.size();
            tk.push_back({fidx, i, j, k | (type << 16)});
            const int kw = std::min(NB, f.nn - NB * k);
            const int m = rows(i), n = type == DAG_POTRF ? kw : (type == DAG_TRSM ? kw : rows(j));
            dur.push_back(dag_cost(type, m, n, kw));
            double flops = 0.0;
            if (type == DAG_POTRF) flops = (double)kw * kw * kw / 3.0;
            else if (type == DAG_TRSM) flops = (double)m * kw * kw;
            else flops = (i == j ? (double)m * (m + 1) : 2.0 * m * n) * kw;
            fl.push_back(flops);
            for (int d : deps)
                if (d >= 0 && std::find(dep.begin() + dptr.back(), dep.end(), d) == dep.end())
                    dep.push_back(d);
            dptr.push_back((int)dep.size());
            return t;
        };
        for (int k = 0; k < F.P; ++k) {
            at(k, k) = add(DAG_POTRF, k, k, k, {at(k, k)});
            for (int i = k + 1; i < R; ++i) at(i, k) = add(DAG_TRSM, i, k, k, {at(i, k), at(k, k)});
            for (int j = k + 1; j < R; ++j)
                for (int i = j; i < R; ++i)
                    at(i, j) = add(DAG_UPDATE, i, j, k, {at(i, j), at(i, k), at(j, k)});
        }
    }
    const int T = (int)tk.size();
    if (T == 0) return;
    // longest path to the end (tasks are created in a topological order)
    std::vector<int> sptr(T + 1, 0), succ(dep.size());
    for (int d : dep) ++sptr[d + 1];
    for (int t = 0; t < T; ++t) sptr[t + 1] += sptr[t];
    {
        std::vector<int> fill(sptr.begin(), sptr.end() - 1);
        for (int t = 0; t < T; ++t)
            for (int e = dptr[t]; e < dptr[t + 1]; ++e) succ[fill[dep[e]]++] = t;
    }
    std::vector<double> bl(T);
    for (int t = T - 1; t >= 0; --t) {
        double m = 0.0;
        for (int e = sptr[t]; e < sptr[t + 1]; ++e) m = std::max(m, bl[succ[e]]);
        bl[t] = dur[t] + m;
    }
    // list schedule on W workers
    const size_t lds = DAG_LDS_BYTES;
    const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(8, P->lds_per_cu / lds));
    const int W = std::min(T, P->ncu * per_cu);
    std::vector<int> indeg(T);
    for (int t = 0; t < T; ++t) indeg[t] = dptr[t + 1] - dptr[t];
    auto cmp = [&](int a, int b) { return bl[a] < bl[b] || (bl[a] == bl[b] && a > b); };
    std::priority_queue<int, std::vector<int>, decltype(cmp)> ready(cmp);
    typedef std::pair<double, int> Ev;
    std::priority_queue<Ev, std::vector<Ev>, std::greater<Ev>> running;
    for (int t = 0; t < T; ++t)
        if (!indeg[t]) ready.push(t);
    std::vector<DagTask> order;
    order.reserve(T);
    double now = 0.0, flops = 0.0;
    int idle = W;
    while ((int)order.size() < T) {
        while (idle > 0 && !ready.empty()) {
            const int t = ready.top();
            ready.pop();
            order.push_back(tk[t]);
            flops += fl[t];
            running.push({now + dur[t], t});
            --idle;
        }
        if (running.empty()) break;  // (cannot happen: the DAG is acyclic)
        const Ev e = running.top();
        running.pop();
        now = e.first;
        ++idle;
        for (int q = sptr[e.second]; q < sptr[e.second + 1]; ++q)
            if (--indeg[succ[q]] == 0) ready.push(succ[q]);
    }
    // the queue order runs sequentially: every wait is met by a task listed
    // before it (what makes the launch's waits end)
    {
        std::vector<int> c(P->dagcnt_len, 0);
        for (const DagTask& t : order) {
            const DagFront& F = fr[t.f];
            const int k = t.kt & 0xffff, type = t.kt >> 16;
            auto cn = [&](int i, int j) -> int& { return c[F.cbase + (size_t)i * (i + 1) / 2 + j]; };
            bool ok = cn(t.i, t.j) == k;
            if (type == DAG_TRSM) ok = ok && cn(k, k) == k + 1;
            if (type == DAG_UPDATE) ok = ok && cn(t.i, k) == k + 1 && cn(t.j, k) == k + 1;
            if (!ok) { P->dag_bad = true; break; }
            ++cn(t.i, t.j);
        }
    }
    const int qidx = (int)P->dagcnt_len++;
    ListBuf& lb = P->pg->lists;
    std::vector<int4> hdr{make_int4(T, (int)fr.size(), qidx, (int)P->dag_tasks)};
    P->dag_tasks += T;
    const i64 off = lb.add(hdr);  // fronts and tasks follow the header contiguously
    const size_t fb = fr.size() * sizeof(DagFront), tb = order.size() * sizeof(DagTask);
    lb.bytes.resize(off + sizeof(int4) + fb + tb);
    std::memcpy(lb.bytes.data() + off + sizeof(int4), fr.data(), fb);
    std::memcpy(lb.bytes.data() + off + sizeof(int4) + fb, order.data(), tb);
    Launch L{K_CHOL_DAG, W, 256, 0, off, T};
    L.flops = flops;
    seq.push_back(L);
}

// The factor's stored convention B21 = L21 L11^{-1} for the large fronts in
// `large`, on side stream 2.  The diagonal blocks were staged in the inverse
// area by k_panel_factor (writing them into blkval in that launch would race
// with workgroups still reading the unfactored block): one launch copies them
// into blkval and replaces the staged copies by their inverses.  Then groups
// of NB2/NB column blocks from the right; inside a group block by block
// (X_b = R_b Dinv_b, update of the group's remaining columns), then one update
// of everything left of the group with K = the group width.  A front's
// conversion reads and writes only its own block and inverse area, and nothing
// later in the factorization reads them (parents read only update matrices),
// so the fronts of many levels convert together, beside the levels above.
void build_chol_convert(chp_plan* P, const std::vector<int>& large, std::vector<Launch>& seq, bool dag) {
    if (large.empty()) return;
    const size_t s0 = seq.size();
    const int nb2 = P->nb2;
    P->pg->split_side = true;
    std::vector<DiagTask> inv;
    for (int k : large) {
        const FrontInfo& f = P->fi[k];
        const int nj = f.nn + f.na;
        for (int b = 0; b * NB < f.nn; ++b) {
            DiagTask t;
            std::memset(&t, 0, sizeof(t));
            t.off = P->h_dinvptr[k] + (i64)b * NB * NB;
            t.outoff = t.off;
            t.ld = NB; t.ldo = NB; t.nb = std::min(NB, f.nn - b * NB); t.key = k; t.col = b * NB;
            t.buf = 3; t.obuf = 3; t.mode = 2;
            t.off2 = f.blkptr + (i64)b * NB * (1 + (i64)nj);
            t.ld2 = nj; t.buf2 = 0;
            inv.push_back(t);
        }
    }
    if (!dag && !inv.empty())
        seq.push_back({K_DIAG, (int)inv.size(), 256, 0, P->pg->lists.add(inv), 0});
    const int GB = nb2 / NB;
    int maxgrp = 0;
    for (int k : large)
        if (P->fi[k].na > 0) maxgrp = std::max(maxgrp, ((P->fi[k].nn + NB - 1) / NB + GB - 1) / GB);
    for (int sg = 0; sg < maxgrp; ++sg) {
        for (int sb = 0; sb < GB; ++sb) {
            std::vector<GemmTile> pt, it;
            for (int k : large) {
                const FrontInfo& f = P->fi[k];
                int nn = f.nn, na = f.na, nj = nn + na;
                int nbk = (nn + NB - 1) / NB, ng = (nbk + GB - 1) / GB;
                if (na == 0 || sg >= ng) continue;
                int g = ng - 1 - sg;
                int gb0 = g * GB, gb1 = std::min(gb0 + GB, nbk);
                int b = gb1 - 1 - sb;
                if (b < gb0) continue;
                int b0 = b * NB, b1 = std::min(b0 + NB, nn), q0 = gb0 * NB;
                // X_b = R_b L_bb^{-1} in place (inverse column-major, ld NB)
                add_gemm_tiles(P, pt, mk(0, f.blkptr + nn + (i64)b0 * nj, nj, 0,
                                         3, P->h_dinvptr[k] + (i64)b * NB * NB, NB, 0,
                                         0, f.blkptr + nn + (i64)b0 * nj, nj,
                                         na, b1 - b0, b1 - b0, 1.0, 0.0, 0), false);
                // R[:, q0:b0] -= X_b L[b0:b1, q0:b0]
                if (b0 > q0)
                    add_gemm_tiles(P, it, mk(0, f.blkptr + nn + (i64)b0 * nj, nj, 0,
                                             0, f.blkptr + b0 + (i64)q0 * nj, nj, 0,
                                             0, f.blkptr + nn + (i64)q0 * nj, nj,
                                             na, b0 - q0, b1 - b0, -1.0, 1.0, 0), false);
            }
            push_gemm(P, seq, pt);
            push_gemm(P, seq, it);
        }
        std::vector<GemmTile> gt;
        for (int k : large) {
            const FrontInfo& f = P->fi[k];
            int nn = f.nn, na = f.na, nj = nn + na;
            int nbk = (nn + NB - 1) / NB, ng = (nbk + GB - 1) / GB;
            if (na == 0 || sg >= ng) continue;
            int g = ng - 1 - sg;
            int q0 = g * GB * NB, q1 = std::min((g * GB + GB) * NB, nn);
            // R[:, 0:q0] -= X[:, q0:q1] L[q0:q1, 0:q0]
            if (q0 > 0)
                add_gemm_tiles(P, gt, mk(0, f.blkptr + nn + (i64)q0 * nj, nj, 0,
                                         0, f.blkptr + q0, nj, 0,
                                         0, f.blkptr + nn, nj, na, q0, q1 - q0, -1.0, 1.0, 0), false);
        }
        push_gemm(P, seq, gt);
    }
    P->pg->split_side = false;
    for (size_t i = s0; i < seq.size(); ++i) seq[i].side = 2;
}

// Cholesky launches for one level (cholesky.c:26-82, all fronts of the level).
// outer panels whose first panel launch has at most this many workgroups
// are factored left-looking inside the outer panel
constexpr int LEFT_LOOKING_MAX_WG = 1024;

void build_chol_level(chp_plan* P, const std::vector<int>& fr, std::vector<Launch>& seq, bool dag,
                      std::vector<int>& convert) {
    std::vector<int> small, large;
    size_t lds = 0;
    for (int k : fr) {
        const FrontInfo& f = P->fi[k];
        if (chol_is_small(P, f)) { small.push_back(k); lds = std::max(lds, chol_small_lds(f)); }
        else large.push_back(k);
    }
    // small fronts (LDS kernel) and large fronts of one level write disjoint
    // front and update storage: the small-front launch runs on the side stream
    const bool fork = !small.empty() && !large.empty();
    if (fork) seq.push_back({K_FORK, 0, 0, 0, 0, 1});
    if (!small.empty()) {
        Launch Ls = small_launch(P, K_CHOL_SMALL, small, lds);
        Ls.side = fork ? 1 : 0;
        seq.push_back(Ls);
    }
    if (large.empty()) return;
    if (dag) P->n_large_chol += (int)large.size();
    int net = 0;
    const std::vector<int4> et = extadd_tiles(P, large, &net);
    seq.push_back({K_EXTADD, net, 256, 0, P->pg->lists.add(et), 1});
    // Two-level blocking: 64-wide inner panels (each launch first applies the
    // outer panel's earlier inner panels to its own columns, left-looking,
    // then factors the diagonal block and solves the panel), then one
    // trailing update with K = the outer panel width for everything right of
    // the outer panel.
    int maxnn = 0;
    for (int k : large) maxnn = std::max(maxnn, P->fi[k].nn);
    const int nb2 = P->nb2;
    if (dag)  // one persistent launch for the whole level (diagonal blocks and
              // their inverses included)
        build_chol_dag(P, large, seq);
    else
    for (int p0 = 0; p0 < maxnn; p0 += nb2) {
        // one choice per outer panel (its inner panels must not apply the
        // same update twice), from the first inner panel's workgroup count
        int nwg = 0;
        for (int k : large) {
            const FrontInfo& f = P->fi[k];
            if (f.nn > p0) nwg += 1 + (f.nn + f.na - std::min(p0 + NB, f.nn) + NB - 1) / NB;
        }
        const bool left = nwg <= LEFT_LOOKING_MAX_WG;
        for (int c0 = p0; c0 < std::min(p0 + nb2, maxnn); c0 += NB) {
            // diagonal block and the panel below it, one launch: a workgroup
            // per 64 panel rows (plus one writing the diagonal block).  When
            // the outer panel's launches have at most LEFT_LOOKING_MAX_WG
            // workgroups, the update from its earlier inner panels (columns
            // p0 .. c0 - 1) comes first in the same launch (left-looking: a
            // GEMM launch between the inner panels costs about 10 us of the
            // critical path); with more, a GEMM launch after each inner panel
            // updates the outer panel's remaining columns, since every
            // workgroup of a left-looking launch forms the diagonal block's
            // update itself.  Measured (40^3 cholesky + projected inverse /
            // 512 batched grids' cholesky): right-looking everywhere 14.47 /
            // 8.63 ms, left-looking up to 1024 workgroups 14.27 / 8.22 ms,
            // left-looking everywhere 14.30 / 8.25 ms.
            std::vector<int4> pf;
            std::vector<GemmTile> it;
            for (int k : large) {
                const FrontInfo& f = P->fi[k];
                int nn = f.nn, nj = nn + f.na;
                if (nn <= c0) continue;
                int c1 = std::min(c0 + NB, nn), p1 = std::min(p0 + nb2, nn);
                const int kp = left ? p0 : c0;
                pf.push_back(make_int4(k, c0, -1, kp));
                for (int r0 = c1; r0 < nj; r0 += NB) pf.push_back(make_int4(k, c0, r0, kp));
                // right-looking: F[c1:, c1:p1] -= L[c1:, c0:c1] L[c1:p1, c0:c1]^T
                if (!left && c1 < p1)
                    add_gemm_tiles(P, it, mk(0, f.blkptr + c1 + (i64)c0 * nj, nj, 0,
                                             0, f.blkptr + c1 + (i64)c0 * nj, nj, 1,
                                             0, f.blkptr + c1 + (i64)c1 * nj, nj,
                                             nj - c1, p1 - c1, c1 - c0, -1.0, 1.0, 1), true);
            }
            seq.push_back({K_PANEL_FACTOR, (int)pf.size(), 512, 0, P->pg->lists.add(pf), 0});
            push_gemm(P, seq, it, 1);
        }
        // trailing update right of the outer panel, K = p1 - p0.  (Splitting
        // off the next outer panel's columns and running the rest on a side
        // stream measured slower: the rest's workgroups occupy every CU, so
        // the next panels wait for CUs instead of overlapping it.)
        std::vector<GemmTile> gt;
        for (int k : large) {
            const FrontInfo& f = P->fi[k];
            int nn = f.nn, na = f.na, nj = nn + na;
            if (nn <= p0) continue;
            int p1 = std::min(p0 + nb2, nn);
            if (p1 < nn)
                add_gemm_tiles(P, gt, mk(0, f.blkptr + p1 + (i64)p0 * nj, nj, 0,
                                         0, f.blkptr + p1 + (i64)p0 * nj, nj, 1,
                                         0, f.blkptr + p1 + (i64)p1 * nj, nj,
                                         nj - p1, nn - p1, p1 - p0, -1.0, 1.0, 1), true);
            if (na > 0)
                add_gemm_tiles(P, gt, mk(0, f.blkptr + nn + (i64)p0 * nj, nj, 0,
                                         0, f.blkptr + nn + (i64)p0 * nj, nj, 1,
                                         1, f.updptr, na, na, na, p1 - p0, -1.0, 1.0, 1), true);
        }
        push_gemm(P, seq, gt, 1);
    }
    convert.insert(convert.end(), large.begin(), large.end());
    if (fork) seq.push_back({K_JOIN, 0, 0, 0, 0, 1});
}

// Projected-inverse launches for one level (projected_inverse.c:25-73).
// X = L11^{-1} of a large front reads only the factor, so it is computed on
// side stream 2, which runs ahead through the levels; every large front has
// its own X at h_xptr[k] in [0, wbase).  W (inversion temporary, side stream
// 2 only) is per level from wbase on, T = V_k B21 (main stream only) per level
// from tbase on: each stream reuses its own region level after level.
void build_projinv_level(chp_plan* P, const std::vector<int>& fr, i64 wbase, i64 tbase,
                         i64& xcur, i64& scratch_need) {
    // V_c of the children of this level's fronts, extracted once the fronts
    // are finished (end of the level): an update slot is touched only at its
    // front's level and its parent's, which lets the arena reuse slots
    const std::vector<int4> xt = extract_tiles(P, children_of(P, fr));
    std::vector<int> small, large;
    size_t lds = 0;
    for (int k : fr) {
        const FrontInfo& f = P->fi[k];
        if (projinv_is_small(P, f)) { small.push_back(k); lds = std::max(lds, projinv_small_lds(f)); }
        else large.push_back(k);
    }
    if (!large.empty()) {
        P->n_large_projinv += (int)large.size();
        i64 woff_end = wbase, toff_end = tbase;
        std::vector<i64> woff(large.size());
        for (size_t q = 0; q < large.size(); ++q) {
            const FrontInfo& f = P->fi[large[q]];
            P->h_xptr[large[q]] = xcur;
            xcur += ((i64)f.nn * f.nn + 31) / 32 * 32;
            woff[q] = woff_end; woff_end += ((i64)f.nn * f.nn + 31) / 32 * 32;
        }
        for (size_t q = 0; q < large.size(); ++q) {
            const FrontInfo& f = P->fi[large[q]];
            P->h_tptr[large[q]] = toff_end;
            toff_end += ((i64)f.na * f.nn + 31) / 32 * 32;
        }
        scratch_need = std::max(scratch_need, toff_end);  // W ends at or below tbase
        // side stream 2: diagonal block inverses, then recursive halving of
        // the block range [a, b) at m:
        //   X[m:b, a:m] = -X[m:b, m:b] (L[m:b, a:m] X[a:m, a:m]);
        // all nodes of one height (over all large fronts) form one launch pair.
        const size_t s0 = P->projinv.size();
        P->pg->split_side = true;
        std::vector<int2> bt;
        for (int k : large) {
            int nbk = (P->fi[k].nn + NB - 1) / NB;
            for (int b = 0; b < nbk; ++b) bt.push_back(make_int2(k, b));
        }
        P->projinv.push_back({K_TRTRI_DIAG, (int)bt.size(), 256, 0, P->pg->lists.add(bt), 0});
        std::vector<std::vector<std::array<int, 4>>> byh;  // (q, a, m, b)
        std::function<int(int, int, int)> tree = [&](int q, int a, int b) -> int {
            if (b - a <= 1) return 0;
            int m = (a + b) / 2;
            int h = 1 + std::max(tree(q, a, m), tree(q, m, b));
            if ((int)byh.size() < h) byh.resize(h);
            byh[h - 1].push_back({q, a, m, b});
            return h;
        };
        for (size_t q = 0; q < large.size(); ++q)
            tree((int)q, 0, (P->fi[large[q]].nn + NB - 1) / NB);
        for (auto& level_nodes : byh) {
            std::vector<GemmTile> g1, g2;
            for (auto& nd : level_nodes) {
                int q = nd[0];
                int k = large[q];
                const FrontInfo& f = P->fi[k];
                int nn = f.nn, nj = nn + f.na;
                int r1 = nd[1] * NB, r2 = nd[2] * NB, r3 = std::min(nd[3] * NB, nn);
                int m1 = r2 - r1, m2 = r3 - r2;
                i64 xo = P->h_xptr[k], wo = woff[q];
                // W[r2:r3, r1:r2] = L[r2:r3, r1:r2] X[r1:r2, r1:r2]   (X lower)
                add_gemm_tiles(P, g1, mk(0, f.blkptr + r2 + (i64)r1 * nj, nj, 0,
                                         2, xo + r1 + (i64)r1 * nn, nn, 0,
                                         2, wo + r2 + (i64)r1 * nn, nn,
                                         m2, m1, m1, 1.0, 0.0, 0, 0, 1), false);
                // X[r2:r3, r1:r2] = -X[r2:r3, r2:r3] W[r2:r3, r1:r2]   (X lower)
                add_gemm_tiles(P, g2, mk(2, xo + r2 + (i64)r2 * nn, nn, 0,
                                         2, wo + r2 + (i64)r1 * nn, nn, 0,
                                         2, xo + r2 + (i64)r1 * nn, nn,
                                         m2, m1, m2, -1.0, 0.0, 0, 1, 0), false);
            }
            push_gemm(P, P->projinv, g1);
            push_gemm(P, P->projinv, g2);
        }
        P->pg->split_side = false;
        for (size_t i = s0; i < P->projinv.size(); ++i) P->projinv[i].side = 2;
    }
    // side stream 1: the small fronts (their V_k were extracted at the end of
    // the parents' level), beside the large fronts
    const bool fork = !large.empty() && !small.empty();
    if (fork) P->projinv.push_back({K_FORK, 0, 0, 0, 0, 1});
    if (!small.empty()) {
        Launch Ls = small_launch(P, K_PROJINV_SMALL, small, lds);
        Ls.side = fork ? 1 : 0;
        P->projinv.push_back(Ls);
    }
    auto extract_children = [&]() {
        if (!xt.empty())
            P->projinv.push_back({K_EXTRACT, (int)xt.size(), 256, 0, P->pg->lists.add(xt), 0});
    };
    if (large.empty()) { extract_children(); return; }
    P->projinv.push_back({K_JOIN, 0, 0, 0, 0, 2});
    // F11 = X^T X (lower, strict upper zeroed) and T = V B21
    std::vector<GemmTile> g3, g4;
    for (int k : large) {
        const FrontInfo& f = P->fi[k];
        int nn = f.nn, na = f.na, nj = nn + na;
        i64 xo = P->h_xptr[k], to = P->h_tptr[k];
        add_gemm_tiles(P, g3, mk(2, xo, nn, 1, 2, xo, nn, 0, 0, f.blkptr, nj,
                                 nn, nn, nn, 1.0, 0.0, 2, 2, 1), false);
        if (na > 0) {
            // V_k is stored with both triangles by the extract kernel
            add_gemm_tiles(P, g3, mk(1, f.updptr, na, 0, 0, f.blkptr + nn, nj, 0,
                                     2, to, na, na, nn, na, 1.0, 0.0, 0), false);
            add_gemm_tiles(P, g4, mk(2, to, na, 1, 0, f.blkptr + nn, nj, 0,
                                     0, f.blkptr, nj, nn, nn, na, 1.0, 1.0, 1), true);
        }
    }
    push_gemm(P, P->projinv, g3);
    push_gemm(P, P->projinv, g4);
    if (fork) P->projinv.push_back({K_JOIN, 0, 0, 0, 0, 1});
    // one launch for the children's V and the large fronts' F21 = -T: the
    // children of large fronts read F21 as -T from scratch (tile w = 1), so
    // the negated copy (tiles w = 2) does not have to precede them
    std::vector<int4> xl = xt;
    std::vector<char> is_large(P->fi.size(), 0);
    for (int k : large) is_large[k] = 1;
    for (int4& x : xl)
        if (is_large[P->fi[x.x].parent] && P->fi[P->fi[x.x].parent].na > 0) x.w = 1;
    for (int k : large)
        if (P->fi[k].na > 0)
            for (int c0 = 0; c0 < P->fi[k].nn; c0 += 16) xl.push_back(make_int4(k, c0, 0, 2));
    if (!xl.empty())
        P->projinv.push_back({K_EXTRACT, (int)xl.size(), 256, 0, P->pg->lists.add(xl), 0});
}

// llt launches for one level (llt.c:26-58).  Large fronts: keep the diagonal
// block as stored (A_top), A_bot = B21 L11, F = A A^T on MFMA (lower), then
// extend-add the children into F without clearing F22.
void build_llt_level(chp_plan* P, const std::vector<int>& fr, i64& scratch_need) {
    std::vector<int> small, large;
    size_t lds = 0;
    for (int k : fr) {
        const FrontInfo& f = P->fi[k];
        if (f.nn + f.na <= P->small_max && llt_small_lds(f) <= P->lds_cap) {
            small.push_back(k);
            lds = std::max(lds, llt_small_lds(f));
        } else {
            large.push_back(k);
        }
    }
    if (!small.empty())
        P->llt.push_back(small_launch(P, K_LLT_SMALL, small, lds));
    if (large.empty()) return;
    // scratch A (nj x nn, ld nj) per large front; offsets in h_xptr are per
    // sweep and the levels run one after the other, so they are reused
    i64 off = 0;
    std::vector<i64> aoff(large.size());
    for (size_t q = 0; q < large.size(); ++q) {
        const FrontInfo& f = P->fi[large[q]];
        aoff[q] = off;
        off += (i64)(f.nn + f.na) * f.nn;
        off = (off + 31) / 32 * 32;
    }
    scratch_need = std::max(scratch_need, off);
    std::vector<int2> ct;
    std::vector<GemmTile> g1, g2;
    for (size_t q = 0; q < large.size(); ++q) {
        int k = large[q];
        const FrontInfo& f = P->fi[k];
        int nn = f.nn, na = f.na, nj = nn + na;
        P->h_lltptr[k] = aoff[q];
        for (int c0 = 0; c0 < nn; c0 += 16) ct.push_back(make_int2(k, c0));
        if (na > 0)  // A_bot = B21 tril(L11)
            add_gemm_tiles(P, g1, mk(0, f.blkptr + nn, nj, 0, 0, f.blkptr, nj, 0,
                                     2, aoff[q] + nn, nj, na, nn, nn, 1.0, 0.0, 0, 0, 1), false);
        // F[:, :nn] (lower) and F22 = A A^T
        add_gemm_tiles(P, g2, mk(2, aoff[q], nj, 0, 2, aoff[q], nj, 1,
                                 0, f.blkptr, nj, nj, nn, nn, 1.0, 0.0, 1), true);
        if (na > 0)
            add_gemm_tiles(P, g2, mk(2, aoff[q] + nn, nj, 0, 2, aoff[q] + nn, nj, 1,
                                     1, f.updptr, na, na, na, nn, 1.0, 0.0, 1), true);
    }
    P->llt.push_back({K_COPY_TOP, (int)ct.size(), 256, 0, P->pg->lists.add(ct), 0});
    push_gemm(P, P->llt, g1);
    push_gemm(P, P->llt, g2);
    int net = 0;
    const std::vector<int4> et = extadd_tiles(P, large, &net);
    P->llt.push_back({K_EXTADD, net, 256, 0, P->pg->lists.add(et), 0});
}

// Offsets of the update slots in the arena (fi[k].updptr), and the arena
// length.  Every sweep writes and reads slot k (na_k^2 doubles) only at
// the levels level(k) and level(parent(k)) (cholesky, llt, _Y2K: written
// at k's level, read at the parent's; projected inverse, completion, the
// scaling maps, _M2T: written at the parent's level, read at k's), and a
// level finishes before the next begins.  So a slot is live over the closed
// level interval [level(k), level(parent(k))], and slots whose intervals
// meet get disjoint memory while the others share it: the arena holds the
// largest set of slots live at one level instead of sum(na^2).  Levels are
// visited bottom-up; slots whose interval ended below the level are
// released (coalescing neighbours), then the level's new slots take the
// best-fitting free block or the end of the arena.
i64 place_update_slots(std::vector<FrontInfo>& fi, const std::vector<std::vector<int>>& bylevel) {
    const int H = (int)bylevel.size();
    std::map<i64, i64> by_off;              // free block offset -> length
    std::set<std::pair<i64, i64>> by_len;   // (length, offset) of free blocks
    std::vector<std::vector<int>> release(H + 1);
    i64 top = 0;
    auto slot_len = [&](int k) { return ((i64)fi[k].na * fi[k].na + 31) / 32 * 32; };
    auto free_block = [&](i64 off, i64 len) {
        auto nx = by_off.lower_bound(off);
        if (nx != by_off.end() && off + len == nx->first) {  // merge with the next block
            by_len.erase({nx->second, nx->first});
            len += nx->second;
            nx = by_off.erase(nx);
        }
        if (nx != by_off.begin()) {
            auto pv = std::prev(nx);
            if (pv->first + pv->second == off) {  // merge into the previous block
                by_len.erase({pv->second, pv->first});
                off = pv->first;
                len += pv->second;
                by_off.erase(pv);
            }
        }
        if (off + len == top) { top = off; return; }  // the arena's tail shrinks
        by_off[off] = len;
        by_len.insert({len, off});
    };
    for (int L = 0; L < H; ++L) {
        for (int k : release[L]) free_block(fi[k].updptr, slot_len(k));
        for (int k : bylevel[L]) {
            FrontInfo& f = fi[k];
            if (f.na == 0) { f.updptr = 0; continue; }
            const i64 need = slot_len(k);
            auto it = by_len.lower_bound({need, (i64)0});
            if (it != by_len.end()) {
                const i64 off = it->second, len = it->first;
                by_len.erase(it);
                by_off.erase(off);
                f.updptr = off;
                if (len > need) { by_off[off + need] = len - need; by_len.insert({len - need, off + need}); }
            } else {
                f.updptr = top;
                top += need;
            }
            // released once the parent's level is over
            const int end = fi[f.parent].level;
            release[std::min(end + 1, H)].push_back(k);
        }
    }
    i64 len = 0;
    for (const FrontInfo& f : fi) len = std::max(len, f.updptr + (i64)f.na * f.na);
    return len;
}

int alloc_doubles(double** d, i64 len) {
    HIP_CHECK(hipMalloc((void**)d, (size_t)std::max<i64>(len, 1) * sizeof(double)));
    return 0;
}

int alloc_ints(int** d, i64 len) {
    HIP_CHECK(hipMalloc((void**)d, (size_t)std::max<i64>(len, 1) * sizeof(int)));
    return 0;
}

// Profiled launches take their start/stop timestamps from the kernel's own
// dispatch (hipExtLaunchKernelGGL events), not from separate event packets.
#define CHP_LAUNCH(K, LDS, ...)                                                        \
    do {                                                                               \
        if (P->prof)                                                                   \
            hipExtLaunchKernelGGL(K, g, b, LDS, s, P->ev0, P->ev1, 0, __VA_ARGS__);    \
        else                                                                           \
            hipLaunchKernelGGL(K, g, b, LDS, s, __VA_ARGS__);                          \
    } while (0)

BufTab main_tab(chp_plan* P, double* blkval) {
    BufTab t;
    t.p[0] = blkval; t.p[1] = P->d_upd; t.p[2] = P->d_scratch; t.p[3] = P->d_dinv;
    t.p[4] = P->d_splitp; t.p[5] = nullptr; t.p[6] = nullptr; t.p[7] = nullptr;
    t.gate = nullptr;
    return t;
}

int sync_program(Program& pg) {
    if (pg.lists.bytes.size() != pg.up_lists || !pg.d_lists) {
        if (pg.d_lists) HIP_CHECK(hipFree(pg.d_lists));
        pg.d_lists = nullptr;
        if (upload(&pg.d_lists, pg.lists.bytes)) return -2;
        pg.up_lists = pg.lists.bytes.size();
    }
    if (pg.tasks.size() != pg.up_tasks || !pg.d_tasks) {
        if (pg.d_tasks) HIP_CHECK(hipFree(pg.d_tasks));
        pg.d_tasks = nullptr;
        if (upload(&pg.d_tasks, pg.tasks)) return -2;
        pg.up_tasks = pg.tasks.size();
    }
    return 0;
}

int ensure_side_stream(chp_plan* P) {
    for (int i = 1; i < chp_plan::NSIDE; ++i) {
        if (P->sstream[i]) continue;
        HIP_CHECK(hipStreamCreateWithFlags(&P->sstream[i], hipStreamNonBlocking));
        HIP_CHECK(hipEventCreateWithFlags(&P->sev_fork[i], hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&P->sev_join[i], hipEventDisableTiming));
    }
    return 0;
}

int run_seq(chp_plan* P, const Program& pg, const std::vector<Launch>& seq, const BufTab& bt,
            hipStream_t s0) {
    const bool two = !P->prof;  // profiling serializes everything on one stream
    if (two && ensure_side_stream(P)) return -2;
    for (const Launch& L : seq) {
        if (L.kid == K_FORK || L.kid == K_JOIN) {
            if (!two) continue;
            const int i = L.iarg;
            if (L.kid == K_FORK) {
                HIP_CHECK(hipEventRecord(P->sev_fork[i], s0));
                HIP_CHECK(hipStreamWaitEvent(P->sstream[i], P->sev_fork[i], 0));
            } else {
                HIP_CHECK(hipEventRecord(P->sev_join[i], P->sstream[i]));
                HIP_CHECK(hipStreamWaitEvent(s0, P->sev_join[i], 0));
            }
            continue;
        }
        const hipStream_t s = (two && L.side) ? P->sstream[L.side] : s0;
        const void* list = pg.d_lists + L.list;
        const FrontInfo* fi = L.fisel ? P->d_fi_full : P->d_fi;
        double* blkval = bt.p[L.bsel];
        dim3 g(L.grid, L.gy), b(L.block);
        switch (L.kid) {
        case K_CHOL_SMALL:
            CHP_LAUNCH(k_chol_small, L.lds, (const int*)list, L.iarg, fi, blkval,
                       P->d_upd, P->d_relidx, P->d_chidx, P->d_status);
            break;
        case K_EXTADD:
            CHP_LAUNCH(k_extadd_tiles, 0, (const int4*)list, fi, blkval,
                       bt.p[1], P->d_relidx, P->d_chidx, L.iarg, L.bstride, L.ustride);
            break;
        case K_COPY_TOP:
            CHP_LAUNCH(k_copy_top, 0, (const int2*)list, fi, blkval,
                       P->d_scratch, P->d_lltptr);
            break;
        case K_POTRF_DIAG:
            CHP_LAUNCH(k_potrf_diag, 0, (const int*)list, L.iarg, fi,
                       blkval, P->d_dinv, P->d_dinvptr, P->d_status);
            break;
        case K_CHOL_DAG:
            CHP_LAUNCH(k_chol_dag, 0, (const int4*)list, blkval, bt.p[1], bt.p[3], P->d_dagcnt,
                       P->d_status, P->dag_wait_limit, P->d_dagtrace);
            break;
        case K_FILL_INT:
            CHP_LAUNCH(k_fill_int, 0, P->d_dagcnt, (i64)L.iarg, 0);
            break;
        case K_PANEL_FACTOR:
            CHP_LAUNCH(k_panel_factor, 0, (const int4*)list, fi, blkval, P->d_dinv,
                       P->d_dinvptr, P->d_status);
            break;
        case K_GEMM:
            if (L.iarg == 1 && P->schur_mode) {  // low-precision Schur updates (256 threads)
                b = dim3(256);
                if (P->schur_mode == 1)  // FP32 arithmetic
                    CHP_LAUNCH(k_gemm_tiles32, 0, (const GemmTile*)list, pg.d_tasks, bt);
                else                     // BF16x3 split precision
                    CHP_LAUNCH(k_gemm_tiles_bf3, 0, (const GemmTile*)list, pg.d_tasks, bt);
            } else if (L.block == 512) {
                CHP_LAUNCH(k_gemm_tiles_lat, 0, (const GemmTile*)list, pg.d_tasks, bt);
            } else {
                CHP_LAUNCH(k_gemm_tiles, 0, (const GemmTile*)list, pg.d_tasks, bt);
            }
            break;
        case K_SPLITK_REDUCE:
            CHP_LAUNCH(k_splitk_reduce, 0, (const GemmTile*)list, pg.d_tasks, bt);
            break;
        case K_GEMM_NARROW:
            CHP_LAUNCH(k_gemm_narrow, 0, (const GemmTile*)list, pg.d_tasks, bt);
            break;
        case K_GEMM_SMALL:
            CHP_LAUNCH(k_gemm_small, 0, (const int*)list, L.iarg, pg.d_tasks, bt);
            break;
        case K_EXTRACT:
            CHP_LAUNCH(k_extract, 0, (const int4*)list, fi, blkval,
                       bt.p[1], P->d_relidx, L.bstride, L.ustride, P->d_scratch, P->d_tptr);
            break;
        case K_PROJINV_SMALL:
            CHP_LAUNCH(k_projinv_small, L.lds, (const int*)list, L.iarg, fi,
                       blkval, P->d_upd, P->d_status);
            break;
        case K_PROJINV_CHAIN:
            CHP_LAUNCH(k_projinv_chain, 0, (const ChainFront*)list, L.iarg, fi, blkval, P->d_upd,
                       P->d_relidx, P->d_chidx, P->d_status);
            break;
        case K_CHOL_CHAIN:
            CHP_LAUNCH(k_chol_chain, 0, (const ChainFront*)list, L.iarg, fi, blkval, P->d_upd,
                       P->d_relidx, P->d_chidx, P->d_status);
            break;
        case K_UF_SETUP:
            CHP_LAUNCH(k_uf_setup, 0, (const char*)list, L.iarg, bt, P->d_relidx);
            break;
        case K_UF_STEP:
            CHP_LAUNCH(k_uf_step, 0, (const UFBlock*)list, bt, P->d_relidx);
            break;
        case K_LLT_CHAIN:
            CHP_LAUNCH(k_llt_chain, 0, (const ChainFront*)list, L.iarg, fi, blkval, P->d_upd,
                       P->d_relidx, P->d_chidx);
            break;
        case K_TRTRI_DIAG:
            CHP_LAUNCH(k_trtri_diag, 0, (const int2*)list, fi, blkval,
                       P->d_dinv, P->d_dinvptr, P->d_scratch, P->d_xptr, P->d_status);
            break;
        case K_LLT_SMALL:
            CHP_LAUNCH(k_llt_small, L.lds, (const int*)list, L.iarg, fi, blkval,
                       P->d_upd, P->d_relidx, P->d_chidx);
            break;
        case K_DIAG:
            CHP_LAUNCH(k_diag_tasks, 0, (const DiagTask*)list, bt, P->d_status);
            break;
        case K_MATOP:
            CHP_LAUNCH(k_matop, 0, (const MatOp*)list, bt);
            break;
        case K_COMP_CHAIN:
            CHP_LAUNCH(k_completion_chain, 0, (const ChainItem*)list, L.iarg, fi, bt,
                       P->d_relidx, P->d_status);
            break;
        case K_UPDATE_FACTOR:  // the 64-thread launches hold the small tasks
            if (L.block == 64)
                CHP_LAUNCH(k_update_factor_small, 0, (const UFTask*)list, bt, P->d_relidx,
                           P->d_status);
            else
                CHP_LAUNCH(k_update_factor, 0, (const UFTask*)list, bt, P->d_relidx, P->d_status);
            break;
        default:
            chp_set_error("internal: unexpected launch kind");
            return -3;
        }
        if (P->prof) {
            HIP_CHECK(hipEventSynchronize(P->ev1));
            float ms = 0.f;
            HIP_CHECK(hipEventElapsedTime(&ms, P->ev0, P->ev1));
            P->prof_ms[L.kid] += ms;
            P->prof_cnt[L.kid] += 1;
            P->prof_flops[L.kid] += L.flops;
            P->prof_rows.insert(P->prof_rows.end(), {(double)L.kid, (double)L.grid, L.flops, (double)ms});
        }
    }
    HIP_CHECK(hipGetLastError());
    return 0;
}
#undef CHP_LAUNCH

// Runs `seq` (preceded by a status reset) on stream s, eagerly or through a
// cached HIP graph: with graphs enabled, the first calls time both ways and
// the faster one is kept (chp_plan::Cached).
int run_sweep(chp_plan* P, const std::vector<Launch>& seq, chp_plan::Cached& G, double* blkval,
              hipStream_t s, bool reset_status) {
    const BufTab bt = main_tab(P, blkval);
    if (!P->use_graphs || P->prof) {
        if (reset_status)
            hipLaunchKernelGGL(k_status_init, dim3(1), dim3(1), 0, s, P->d_status);
        return run_seq(P, P->main, seq, bt, s);
    }
    if (!P->gstream) {
        HIP_CHECK(hipStreamCreateWithFlags(&P->gstream, hipStreamNonBlocking));
        HIP_CHECK(hipEventCreateWithFlags(&P->gev_in, hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&P->gev_out, hipEventDisableTiming));
    }
    if (ensure_side_stream(P)) return -2;  // created outside any capture
    if (!G.e0) {
        HIP_CHECK(hipEventCreate(&G.e0));
        HIP_CHECK(hipEventCreate(&G.e1));
    }
    if (G.pending) {  // the timed call before this one has been issued: read it
        HIP_CHECK(hipEventSynchronize(G.e1));
        float ms = 0.f;
        HIP_CHECK(hipEventElapsedTime(&ms, G.e0, G.e1));
        (G.pending == 1 ? G.t_eager : G.t_graph) = ms;
        G.pending = 0;
    }
    int mode = G.mode;
    if (!mode) {
        if (G.calls < 2) mode = 1;
        else if (G.calls < 4) mode = 2;
        else mode = G.mode = (G.t_graph <= G.t_eager ? 2 : 1);
    }
    const bool timed = !G.mode && (G.calls == 1 || G.calls == 3);
    ++G.calls;
    if (timed) HIP_CHECK(hipEventRecord(G.e0, s));
    if (mode == 1) {
        if (reset_status)
            hipLaunchKernelGGL(k_status_init, dim3(1), dim3(1), 0, s, P->d_status);
        int rc = run_seq(P, P->main, seq, bt, s);
        if (rc) return rc;
    } else {
        if (G.key != (const void*)blkval || !G.exec) {
            if (G.exec) { HIP_CHECK(hipGraphExecDestroy(G.exec)); G.exec = nullptr; }
            hipGraph_t graph = nullptr;
            HIP_CHECK(hipStreamBeginCapture(P->gstream, hipStreamCaptureModeThreadLocal));
            if (reset_status)
                hipLaunchKernelGGL(k_status_init, dim3(1), dim3(1), 0, P->gstream, P->d_status);
            int rc = run_seq(P, P->main, seq, bt, P->gstream);
            hipError_t e = hipStreamEndCapture(P->gstream, &graph);
            if (rc) { if (graph) (void)hipGraphDestroy(graph); return rc; }
            if (e != hipSuccess) { chp_set_error(std::string("graph capture: ") + hipGetErrorString(e)); return -2; }
            e = hipGraphInstantiate(&G.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) { chp_set_error(std::string("graph instantiate: ") + hipGetErrorString(e)); return -2; }
            G.key = blkval;
        }
        HIP_CHECK(hipEventRecord(P->gev_in, s));
        HIP_CHECK(hipStreamWaitEvent(P->gstream, P->gev_in, 0));
        HIP_CHECK(hipGraphLaunch(G.exec, P->gstream));
        HIP_CHECK(hipEventRecord(P->gev_out, P->gstream));
        HIP_CHECK(hipStreamWaitEvent(s, P->gev_out, 0));
    }
    if (timed) {
        HIP_CHECK(hipEventRecord(G.e1, s));
        G.pending = mode;
    }
    return 0;
}

int read_status(chp_plan* P, hipStream_t s, i64* info) {
    unsigned long long hw[2] = {CHP_STATUS_OK, 0};
    HIP_CHECK(hipMemcpyAsync(hw, P->d_status, sizeof(hw), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (hw[1]) {
        chp_set_error("internal: a dependency wait of the large-front factorization timed out");
        return -3;
    }
    const unsigned long long h = hw[0];
    if (h == CHP_STATUS_OK) {
        if (info) { info[0] = -1; info[1] = 0; }
        return 0;
    }
    if (info) { info[0] = (i64)(h >> 32); info[1] = (i64)(h & 0xffffffffULL); }
    return 1;
}

}  // namespace chpi

using namespace chpi;

extern "C" {

const char* chp_last_error(void) { return g_err.c_str(); }

int chp_version(void) { return 1; }

// See include/chompack_amd.h.  Host only: no device call.
i64 chp_update_layout(i64 nsn, const i64* snpost, const i64* relptr, const i64* chptr,
                      const i64* chidx, i64* updptr, i64* level_out) {
    std::vector<FrontInfo> fi(nsn);
    std::vector<int> level(nsn, 0);
    for (i64 k = 0; k < nsn; ++k) {
        std::memset(&fi[k], 0, sizeof(FrontInfo));
        fi[k].parent = (int)k;
        fi[k].na = (int)(relptr[k + 1] - relptr[k]);
    }
    for (i64 k = 0; k < nsn; ++k)
        for (i64 l = chptr[k]; l < chptr[k + 1]; ++l) {
            if (chidx[l] < 0 || chidx[l] >= nsn) { chp_set_error("invalid chidx"); return -1; }
            fi[chidx[l]].parent = (int)k;
        }
    for (i64 t = 0; t < nsn; ++t) {
        const i64 k = snpost[t];
        if (k < 0 || k >= nsn) { chp_set_error("invalid snpost"); return -1; }
        for (i64 l = chptr[k]; l < chptr[k + 1]; ++l)
            level[k] = std::max(level[k], level[chidx[l]] + 1);
    }
    int H = 0;
    for (i64 k = 0; k < nsn; ++k) { fi[k].level = level[k]; H = std::max(H, level[k] + 1); }
    std::vector<std::vector<int>> bylevel(H);
    for (i64 t = 0; t < nsn; ++t) bylevel[level[snpost[t]]].push_back((int)snpost[t]);
    const i64 len = place_update_slots(fi, bylevel);
    for (i64 k = 0; k < nsn; ++k) {
        if (updptr) updptr[k] = fi[k].updptr;
        if (level_out) level_out[k] = level[k];
    }
    return len;
}

chp_plan* chp_plan_create(i64 n, i64 nsn, const i64* snpost, const i64* snptr,
                          const i64* relptr, const i64* relidx, const i64* chptr,
                          const i64* chidx, const i64* blkptr, const i64* perm,
                          int small_max) {
    const int params[1] = {small_max};
    return chp_plan_create_params(n, nsn, snpost, snptr, relptr, relidx, chptr, chidx, blkptr,
                                  perm, params, 1);
}

chp_plan* chp_plan_create_params(i64 n, i64 nsn, const i64* snpost, const i64* snptr,
                                 const i64* relptr, const i64* relidx, const i64* chptr,
                                 const i64* chidx, const i64* blkptr, const i64* perm,
                                 const int* params, int nparams) {
    chp_plan* P = new chp_plan();
    auto fail = [&](const std::string& m) -> chp_plan* { chp_set_error(m); delete P; return nullptr; };
    if (hipGetDevice(&P->device) != hipSuccess) return fail("no HIP device");
    P->n = n; P->nsn = nsn;
    auto param = [&](int i) { return (params && i < nparams) ? params[i] : 0; };
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, P->device) == hipSuccess && v > 0)
        P->ncu = v;
    size_t lds_block = 64 * 1024;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, P->device) == hipSuccess && v > 0)
        lds_block = (size_t)v;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, P->device) == hipSuccess && v > 0)
        P->lds_per_cu = (size_t)v;
    else
        P->lds_per_cu = lds_block;
    P->small_max = param(0) > 0 ? param(0) : 64;
    if (param(1) > 0) {
        if (param(1) % NB) return fail("outer panel width must be a multiple of 64");
        P->nb2 = param(1);
    }
    if (param(2) > 0) P->splitk_max = param(2);
    if (param(4) > 0) P->chol_dag = param(4) == 1;
    P->lds_cap = std::min<size_t>(param(3) > 0 ? (size_t)param(3) : 64 * 1024, lds_block);
    if (P->lds_cap > 64 * 1024) {  // small-front kernels may take more than 64 KB of LDS
        for (const void* k : {(const void*)k_chol_small, (const void*)k_projinv_small,
                              (const void*)k_llt_small})
            if (hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)P->lds_cap) != hipSuccess)
                return fail("the device refuses " + std::to_string(P->lds_cap) +
                            " bytes of LDS for a small-front kernel");
    }
    std::vector<int> parent(nsn), level(nsn, 0);
    for (i64 k = 0; k < nsn; ++k) parent[k] = (int)k;
    for (i64 k = 0; k < nsn; ++k)
        for (i64 l = chptr[k]; l < chptr[k + 1]; ++l) {
            if (chidx[l] < 0 || chidx[l] >= nsn) return fail("invalid chidx");
            parent[chidx[l]] = (int)k;
        }
    for (i64 t = 0; t < nsn; ++t) {
        i64 k = snpost[t];
        if (k < 0 || k >= nsn) return fail("invalid snpost");
        for (i64 l = chptr[k]; l < chptr[k + 1]; ++l)
            level[k] = std::max(level[k], level[chidx[l]] + 1);
    }
    int H = 0;
    for (i64 k = 0; k < nsn; ++k) H = std::max(H, level[k] + 1);
    P->nlevels = H;
    P->bylevel.assign(H, {});
    for (i64 t = 0; t < nsn; ++t) P->bylevel[level[snpost[t]]].push_back((int)snpost[t]);
    P->fi.resize(nsn);
    i64 upd = 0, dinv = 0, wun = 0;
    P->h_dinvptr.assign(nsn, 0);
    P->h_xptr.assign(nsn, 0);
    P->h_tptr.assign(nsn, 0);
    P->h_wunit.assign(nsn, 0);
    P->h_lltptr.assign(nsn, 0);
    for (i64 k = 0; k < nsn; ++k) {
        FrontInfo& f = P->fi[k];
        f.nn = (int)(snptr[k + 1] - snptr[k]);
        f.na = (int)(relptr[k + 1] - relptr[k]);
        if ((blkptr[k + 1] - blkptr[k]) != (i64)f.nn * (f.nn + f.na))
            return fail("blkptr inconsistent with snptr/relptr");
        f.chbeg = (int)chptr[k];
        f.chend = (int)chptr[k + 1];
        f.parent = parent[k];
        f.level = level[k];
        f.blkptr = blkptr[k];
        f.updptr = upd;  // one slot per front here; compacted below
        f.relptr = relptr[k];
        upd += (i64)f.na * f.na;
        if (!chol_is_small(P, f) || !projinv_is_small(P, f)) {
            P->h_dinvptr[k] = dinv;
            dinv += (i64)((f.nn + NB - 1) / NB) * NB * NB;
        }
        P->h_wunit[k] = wun;
        wun += f.nn + f.na;
    }
    P->blk_len = blkptr[nsn];
    P->fi_full = P->fi;
    P->upd_full_len = upd;
    P->upd_len = place_update_slots(P->fi, P->bylevel);
    P->dinv_len = dinv;
    P->solve_len = wun;
    // host copies of the index maps (the sweep builders read them)
    P->h_relidx.resize(relptr[nsn]);
    P->h_chidx.resize(chptr[nsn]);
    for (i64 i = 0; i < relptr[nsn]; ++i) P->h_relidx[i] = (int)relidx[i];
    for (i64 i = 0; i < chptr[nsn]; ++i) P->h_chidx[i] = (int)chidx[i];
    auto one_small = [&](int L) {
        if (P->bylevel[L].size() != 1) return false;
        const FrontInfo& f = P->fi[P->bylevel[L][0]];
        return f.nn + f.na <= std::min(P->small_max, CHAIN_NJ);
    };
    auto has_large = [&](int L) {
        for (int k : P->bylevel[L])
            if (!chol_is_small(P, P->fi[k])) return true;
        return false;
    };
    auto build_seq = [&](std::vector<Launch>& seq, bool dag) {
        int nbig = 0;
        for (int L = 0; L < H; ++L) nbig += has_large(L);
        // the conversions of the levels with large fronts run in groups of
        // four levels, forked after the group's last level, and the group
        // before the last level is forked before that level (each fork of
        // side stream 2 from the sweep's stream costs the sweep's next launch
        // about 9 us in a graph replay).  Measured on the 40^3 grid (cholesky
        // + projected inverse) / 512 batched grids (cholesky): a fork per
        // level 13.99 / 8.27 ms; everything but the last level in one group
        // 13.78 / 8.34 ms; groups of four 13.67 / 8.10 ms; every conversion
        // after the last level 14.20 ms.
        const int flush_at = std::max(0, nbig - 1);
        std::vector<int> pending;
        auto flush = [&]() {
            if (pending.empty()) return;
            seq.push_back({K_FORK, 0, 0, 0, 0, 2});
            build_chol_convert(P, pending, seq, dag);
            pending.clear();
        };
        int bi = 0;
        for (int L = 0; L < H;) {
            int L2 = L;
            while (L2 < H && one_small(L2)) ++L2;
            if (L2 - L >= CHAIN_MIN) {
                std::vector<int> run;
                for (int M = L; M < L2; ++M) run.push_back(P->bylevel[M][0]);
                seq.push_back(chain_launch(P, K_CHOL_CHAIN, run, true));
                L = L2;
                continue;
            }
            const bool big = has_large(L);
            if (big && bi == flush_at) flush();
            build_chol_level(P, P->bylevel[L], seq, dag, pending);
            if (big) {
                if (bi >= flush_at || (bi + 1) % 4 == 0) flush();
                ++bi;
            }
            ++L;
        }
        flush();
        seq.push_back({K_JOIN, 0, 0, 0, 0, 2});
    };
    build_seq(P->chol, P->chol_dag);
    P->building_lowprec = true;
    build_seq(P->chol_lp, false);
    P->building_lowprec = false;
    if (P->dag_bad) return fail("internal: tile-DAG task order");
    if (P->dagcnt_len > 0) {  // the DAG launches' counters start from zero
        const int g = (int)std::min<i64>(1024, (P->dagcnt_len + 255) / 256);
        P->chol.insert(P->chol.begin(), Launch{K_FILL_INT, g, 256, 0, 0, (int)P->dagcnt_len});
    }
    // scratch of the projected inverse: X = L11^{-1} of every large front,
    // then the largest level's W, then the largest level's T
    i64 wbase = 0, wmax = 0;
    for (int L = 0; L < H; ++L) {
        i64 w = 0;
        for (int k : P->bylevel[L]) {
            const FrontInfo& f = P->fi[k];
            if (projinv_is_small(P, f)) continue;
            const i64 sq = ((i64)f.nn * f.nn + 31) / 32 * 32;
            wbase += sq;
            w += sq;
        }
        wmax = std::max(wmax, w);
    }
    i64 scratch = wbase + wmax, xcur = 0;
    // side stream 2 (the inversions) starts from the sweep's stream
    P->projinv.push_back({K_FORK, 0, 0, 0, 0, 2});
    for (int L = H - 1; L >= 0;) {
        int L2 = L;
        while (L2 >= 0 && one_small(L2)) --L2;
        if (L - L2 >= CHAIN_MIN) {
            std::vector<int> run;
            for (int M = L; M > L2; --M) run.push_back(P->bylevel[M][0]);
            P->projinv.push_back(chain_launch(P, K_PROJINV_CHAIN, run, false));
            L = L2;
            continue;
        }
        build_projinv_level(P, P->bylevel[L], wbase, wbase + wmax, xcur, scratch);
        --L;
    }
    P->projinv.push_back({K_JOIN, 0, 0, 0, 0, 2});
    if (xcur != wbase) return fail("projected-inverse scratch layout");
    for (int L = 0; L < H;) {
        int L2 = L;
        while (L2 < H && one_small(L2)) ++L2;
        if (L2 - L >= CHAIN_MIN) {
            std::vector<int> run;
            for (int M = L; M < L2; ++M) run.push_back(P->bylevel[M][0]);
            P->llt.push_back(chain_launch(P, K_LLT_CHAIN, run, true));
            L = L2;
            continue;
        }
        build_llt_level(P, P->bylevel[L], scratch);
        ++L;
    }
    P->scratch_len = scratch;
    // device copies
    std::vector<i64> pv(perm, perm + n), sp(snptr, snptr + nsn + 1);
    std::vector<int> colfront(n), colidx(n);
    for (i64 k = 0; k < nsn; ++k)
        for (i64 c = snptr[k]; c < snptr[k + 1]; ++c) {
            colfront[c] = (int)k;
            colidx[c] = (int)(c - snptr[k]);
        }
    int rc = 0;
    rc |= upload(&P->d_fi, P->fi);
    rc |= upload(&P->d_fi_full, P->fi_full);
    rc |= upload(&P->d_relidx, P->h_relidx);
    rc |= upload(&P->d_chidx, P->h_chidx);
    rc |= upload(&P->d_perm, pv);
    rc |= upload(&P->d_snptr, sp);
    rc |= upload(&P->d_dinvptr, P->h_dinvptr);
    rc |= upload(&P->d_xptr, P->h_xptr);
    rc |= upload(&P->d_tptr, P->h_tptr);
    rc |= upload(&P->d_wunit, P->h_wunit);
    {
        // forward-solve gather map: per slab row, the children's update rows
        // that land there, children in reverse order
        std::vector<int> gptr(P->solve_len + 1, 0);
        for (i64 k = 0; k < nsn; ++k)
            for (int l = P->fi[k].chend - 1; l >= P->fi[k].chbeg; --l) {
                const FrontInfo& g = P->fi[P->h_chidx[l]];
                for (int q = 0; q < g.na; ++q)
                    ++gptr[P->h_wunit[k] + P->h_relidx[g.relptr + q] + 1];
            }
        for (i64 s = 0; s < P->solve_len; ++s) gptr[s + 1] += gptr[s];
        std::vector<GEnt> gent(gptr[P->solve_len]);
        std::vector<int> fill(gptr.begin(), gptr.end() - 1);
        for (i64 k = 0; k < nsn; ++k)
            for (int l = P->fi[k].chend - 1; l >= P->fi[k].chbeg; --l) {
                const int c = P->h_chidx[l];
                const FrontInfo& g = P->fi[c];
                for (int q = 0; q < g.na; ++q) {
                    GEnt e;
                    e.wu = P->h_wunit[c];
                    e.ld = g.nn + g.na;
                    e.row = g.nn + q;
                    gent[fill[P->h_wunit[k] + P->h_relidx[g.relptr + q]]++] = e;
                }
            }
        rc |= upload(&P->d_gptr, gptr);
        rc |= upload(&P->d_gent, gent);
    }
    rc |= upload(&P->d_lltptr, P->h_lltptr);
    rc |= upload(&P->d_colfront, colfront);
    rc |= upload(&P->d_colidx, colidx);
    P->main.place_side_splits();
    rc |= sync_program(P->main);
    rc |= alloc_doubles(&P->d_upd, P->upd_len);
    rc |= alloc_doubles(&P->d_dinv, P->dinv_len);
    rc |= alloc_doubles(&P->d_scratch, P->scratch_len);
    rc |= alloc_doubles(&P->d_partial, n);
    rc |= alloc_doubles(&P->d_splitp, P->main.split_doubles());
    if (!rc && hipMemset(P->d_splitp, 0, P->main.split_doubles() * sizeof(double)) != hipSuccess) rc = -1;
    if (hipMalloc((void**)&P->d_status, 2 * sizeof(unsigned long long)) != hipSuccess) rc = -1;
    if (!rc) rc |= alloc_ints(&P->d_dagcnt, P->dagcnt_len);
    if (rc) {
        std::string e = g_err;
        chp_plan_destroy(P);
        chp_set_error("plan allocation failed: " + e);
        return nullptr;
    }
    return P;
}

void chp_plan_destroy(chp_plan* P) {
    if (!P) return;
    (void)hipDeviceSynchronize();  // no launch of this plan may still be in flight
    if (P->ev0) (void)hipEventDestroy(P->ev0);
    if (P->ev1) (void)hipEventDestroy(P->ev1);
    for (chp_plan::Cached* G : {&P->g_chol, &P->g_projinv, &P->g_llt, &P->g_chol32, &P->g_cholbf}) {
        if (G->exec) (void)hipGraphExecDestroy(G->exec);
        if (G->e0) (void)hipEventDestroy(G->e0);
        if (G->e1) (void)hipEventDestroy(G->e1);
    }
    if (P->gev_in) (void)hipEventDestroy(P->gev_in);
    if (P->gev_out) (void)hipEventDestroy(P->gev_out);
    if (P->gstream) (void)hipStreamDestroy(P->gstream);
    for (int i = 1; i < chp_plan::NSIDE; ++i) {
        if (P->sev_fork[i]) (void)hipEventDestroy(P->sev_fork[i]);
        if (P->sev_join[i]) (void)hipEventDestroy(P->sev_join[i]);
        if (P->sstream[i]) (void)hipStreamDestroy(P->sstream[i]);
    }
    void* ptrs[] = {P->d_fi, P->d_fi_full, P->d_relidx, P->d_chidx, P->d_perm, P->d_snptr, P->d_dinvptr,
                    P->d_xptr, P->d_tptr, P->d_wunit, P->d_gptr, P->d_gent, P->d_lltptr, P->d_colfront, P->d_colidx, P->d_upd,
                    P->d_dinv, P->d_scratch, P->d_partial, P->d_splitp, P->d_work, P->d_status,
                    P->d_dagcnt, P->d_dagtrace, P->main.d_lists, P->main.d_tasks, P->dense.d_lists, P->dense.d_tasks,
                    P->d_dscratch, P->d_ddinv, P->d_dsplitp, P->d_supd, P->d_ustack, P->d_uupd};
    for (SolvePlan* S : P->solves) {
        for (auto& g : S->graphs) if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (S->pg.d_lists) (void)hipFree(S->pg.d_lists);
        if (S->pg.d_tasks) (void)hipFree(S->pg.d_tasks);
        for (double* d : {S->d_splitp, S->d_inv, S->d_w, S->d_tmp, S->d_l11})
            if (d) (void)hipFree(d);
        if (S->d_chk) (void)hipFree(S->d_chk);
        if (S->d_changed) (void)hipFree(S->d_changed);
        delete S;
    }
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete P;
}

// params[0..3] of the plan in use: small_max, nb2, splitk_max, lds_cap.
int chp_plan_params(chp_plan* P, int* out) {
    if (!P || !out) { chp_set_error("null argument"); return -1; }
    out[0] = P->small_max; out[1] = P->nb2; out[2] = P->splitk_max; out[3] = (int)P->lds_cap;
    out[4] = P->chol_dag ? 1 : 2;
    return 0;
}

// out[0..7] = nlevels, n_large_chol, n_large_projinv, n_gemm_tiles,
//             upd_len, dinv_len, scratch_len, n launches (cholesky)
int chp_plan_info(chp_plan* P, i64* out) {
    if (!P || !out) return -1;
    out[0] = P->nlevels; out[1] = P->n_large_chol; out[2] = P->n_large_projinv;
    out[3] = P->n_gemm_tiles; out[4] = P->upd_len; out[5] = P->dinv_len;
    out[6] = P->scratch_len; out[7] = (i64)P->chol.size();
    return 0;
}

// Per-launch HIP-event timing of the sweeps (adds a synchronization after
// every launch; for measurement only).  on != 0 enables and clears counters.
int chp_profile(chp_plan* P, int on) {
    if (!P) return -1;
    if (on && !P->ev0) {
        HIP_CHECK(hipEventCreate(&P->ev0));
        HIP_CHECK(hipEventCreate(&P->ev1));
    }
    P->prof = on != 0;
    for (int i = 0; i < NKIDS; ++i) { P->prof_ms[i] = 0; P->prof_cnt[i] = 0; P->prof_flops[i] = 0; }
    P->prof_rows.clear();
    return 0;
}

// Per-launch rows (kind, workgroups, algorithmic flops, ms) of the profiled
// sweeps; copies at most max rows into out[4*max] and returns the row count.
i64 chp_profile_launches(chp_plan* P, double* out, i64 max) {
    if (!P) return -1;
    i64 rows = (i64)P->prof_rows.size() / 4;
    for (i64 i = 0; i < std::min(rows, max) * 4; ++i) out[i] = P->prof_rows[i];
    return rows;
}

int64_t chp_dag_trace(chp_plan* P, int on) {
    if (!P) { chp_set_error("null argument"); return -1; }
    if (on && !P->d_dagtrace && P->dag_tasks > 0) {
        if (hipMalloc((void**)&P->d_dagtrace, (size_t)P->dag_tasks * 4 * sizeof(i64)) != hipSuccess) {
            chp_set_error("trace allocation failed");
            return -2;
        }
        (void)hipMemset(P->d_dagtrace, 0, (size_t)P->dag_tasks * 4 * sizeof(i64));
    } else if (!on && P->d_dagtrace) {
        (void)hipDeviceSynchronize();
        (void)hipFree(P->d_dagtrace);
        P->d_dagtrace = nullptr;
    }
    // the captured graphs hold the trace pointer: drop them
    for (chp_plan::Cached* G : {&P->g_chol}) {
        if (G->exec) (void)hipGraphExecDestroy(G->exec);
        G->exec = nullptr;
        G->key = nullptr;
    }
    return P->dag_tasks;
}

int chp_set_dag_wait_limit(chp_plan* P, int polls) {
    if (!P || polls < 0) { chp_set_error("invalid argument"); return -1; }
    P->dag_wait_limit = polls;
    for (chp_plan::Cached* G : {&P->g_chol}) {  // the graphs hold the old bound
        if (G->exec) (void)hipGraphExecDestroy(G->exec);
        G->exec = nullptr;
        G->key = nullptr;
    }
    return 0;
}

int64_t chp_dag_trace_read(chp_plan* P, int64_t* out, int64_t max) {
    if (!P || !out) { chp_set_error("null argument"); return -1; }
    if (!P->d_dagtrace) return 0;
    const i64 n = std::min<i64>(max, P->dag_tasks);
    HIP_CHECK(hipDeviceSynchronize());
    HIP_CHECK(hipMemcpy(out, P->d_dagtrace, (size_t)n * 4 * sizeof(i64), hipMemcpyDeviceToHost));
    return n;
}

int chp_profile_read(chp_plan* P, double* ms, i64* cnt, double* flops, int max) {
    if (!P) return -1;
    for (int i = 0; i < std::min(NKIDS, max); ++i) {
        ms[i] = P->prof_ms[i]; cnt[i] = P->prof_cnt[i]; flops[i] = P->prof_flops[i];
    }
    return NKIDS;
}

// Enable (on != 0, the default) or disable HIP-graph replay of the sweeps.
int chp_set_graphs(chp_plan* P, int on) {
    if (!P) return -1;
    P->use_graphs = on != 0;
    return 0;
}

// Algorithmic GEMM bytes and launch counts: out[0] / out[1] bytes of the
// cholesky / projected-inverse GEMM launches (each task reads A, B and, with
// beta != 0, C once and writes C once), out[2] / out[3] their launch counts.
int chp_plan_gemm_bytes(chp_plan* P, double* out) {
    if (!P || !out) return -1;
    const std::vector<Launch>* seqs[2] = {&P->chol, &P->projinv};
    for (int q = 0; q < 2; ++q) {
        double bytes = 0.0, n = 0.0;
        for (const Launch& L : *seqs[q]) {
            if (L.kid != K_GEMM) continue;
            n += 1.0;
            std::vector<GemmTile> tl(L.grid);
            std::memcpy(tl.data(), P->main.lists.bytes.data() + L.list, L.grid * sizeof(GemmTile));
            int last = -1;
            for (const GemmTile& g : tl) {
                if (g.task == last) continue;
                last = g.task;
                const GemmTask& t = P->main.tasks[g.task];
                bytes += 8.0 * ((double)t.M * t.K + (double)t.K * t.N +
                                (double)t.M * t.N * (t.beta != 0.0 ? 2.0 : 1.0));
            }
        }
        out[q] = bytes;
        out[2 + q] = n;
    }
    return 0;
}

// Algorithmic GEMM flops of one sweep: out[0] cholesky, out[1] projected inverse.
int chp_plan_flops(chp_plan* P, double* out) {
    if (!P) return -1;
    out[0] = out[1] = 0.0;
    for (const Launch& L : P->chol) out[0] += L.flops;
    for (const Launch& L : P->projinv) out[1] += L.flops;
    return 0;
}

int chp_cholesky(chp_plan* P, double* blkval, void* stream, i64* info) {
    if (!P || !blkval) { chp_set_error("null argument"); return -1; }
    hipStream_t s = (hipStream_t)stream;
    int rc = run_sweep(P, P->chol, P->g_chol, blkval, s, true);
    if (rc) return rc;
    return read_status(P, s, info);
}

int chp_cholesky_lowprec(chp_plan* P, double* blkval, int schur, void* stream, i64* info) {
    if (!P || !blkval) { chp_set_error("null argument"); return -1; }
    if (schur != 1 && schur != 2) { chp_set_error("schur must be 1 (FP32) or 2 (BF16x3)"); return -1; }
    hipStream_t s = (hipStream_t)stream;
    P->schur_mode = schur;
    int rc = run_sweep(P, P->chol_lp, schur == 1 ? P->g_chol32 : P->g_cholbf, blkval, s, true);
    P->schur_mode = 0;
    if (rc) return rc;
    return read_status(P, s, info);
}

int chp_cholesky_mixed(chp_plan* P, double* blkval, void* stream, i64* info) {
    return chp_cholesky_lowprec(P, blkval, 1, stream, info);
}

int chp_projected_inverse(chp_plan* P, double* blkval, void* stream, i64* info) {
    if (!P || !blkval) { chp_set_error("null argument"); return -1; }
    hipStream_t s = (hipStream_t)stream;
    int rc = run_sweep(P, P->projinv, P->g_projinv, blkval, s, true);
    if (rc) return rc;
    return read_status(P, s, info);
}

int chp_llt(chp_plan* P, double* blkval, void* stream) {
    if (!P || !blkval) { chp_set_error("null argument"); return -1; }
    int rc = run_sweep(P, P->llt, P->g_llt, blkval, (hipStream_t)stream, false);
    if (rc) return rc;
    HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
    return 0;
}

int chp_dot(chp_plan* P, const double* x, const double* y, double* out, void* stream) {
    if (!P || !x || !y || !out) { chp_set_error("null argument"); return -1; }
    hipStream_t s = (hipStream_t)stream;
    if (P->n == 0) { *out = 0.0; return 0; }
    hipLaunchKernelGGL(k_dot_partial, dim3((unsigned)P->n), dim3(256), 0, s, P->d_colidx,
                       (int)P->n, P->d_fi, P->d_colfront, x, y, P->d_partial);
    HIP_CHECK(hipGetLastError());
    std::vector<double> h(P->n);
    HIP_CHECK(hipMemcpyAsync(h.data(), P->d_partial, P->n * sizeof(double),
                             hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    double v = 0.0;
    for (i64 j = 0; j < P->n; ++j) v += h[j];
    *out = v;
    return 0;
}

int chp_logdet(chp_plan* P, const double* lblkval, double* out, void* stream) {
    if (!P || !lblkval || !out) { chp_set_error("null argument"); return -1; }
    hipStream_t s = (hipStream_t)stream;
    if (P->n == 0) { *out = 0.0; return 0; }
    hipLaunchKernelGGL(k_logdiag, dim3((unsigned)((P->n + 255) / 256)), dim3(256), 0, s,
                       P->d_colidx, (int)P->n, P->d_fi, P->d_colfront, lblkval, P->d_partial);
    HIP_CHECK(hipGetLastError());
    std::vector<double> h(P->n);
    HIP_CHECK(hipMemcpyAsync(h.data(), P->d_partial, P->n * sizeof(double),
                             hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    double v = 0.0;
    for (i64 j = 0; j < P->n; ++j) v += h[j];
    *out = 2.0 * v;
    return std::isfinite(v) ? 0 : 1;
}

int chp_spmv_residual(i64 n, const i64* rowptr, const int* colind, const double* vals,
                      const double* X, int ldx, const double* B, int ldb, double* R, int ldr,
                      int nrhs, double* work, void* stream) {
    if (n <= 0 || nrhs <= 0) return 0;
    if (!rowptr || !colind || !vals || !X || !B || !R) { chp_set_error("null argument"); return -1; }
    if (work && nrhs > 1) {
        // X to row-major in work (n x nrhs), then the row-tiled product
        const dim3 g((unsigned)((n + 63) / 64), (unsigned)((nrhs + 63) / 64));
        hipLaunchKernelGGL(k_to_rowmajor, g, dim3(256), 0, (hipStream_t)stream, n, nrhs, X, ldx,
                           work);
        hipLaunchKernelGGL(k_spmv_residual_rm, g, dim3(256), 0, (hipStream_t)stream, n, rowptr,
                           colind, vals, (const double*)work, nrhs, B, ldb, R, ldr);
        HIP_CHECK(hipGetLastError());
        return 0;
    }
    const i64 total = n * (i64)nrhs;
    const i64 blocks = std::min<i64>((total + 255) / 256, 1 << 20);
    hipLaunchKernelGGL(k_spmv_residual, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       n, rowptr, colind, vals, X, ldx, B, ldb, R, ldr, nrhs);
    HIP_CHECK(hipGetLastError());
    return 0;
}

int chp_copy(double* dst, const double* src, i64 n, void* stream) {
    if (n <= 0) return 0;
    if (!dst || !src || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15)) {
        chp_set_error("chp_copy: null or misaligned (16 bytes) pointer");
        return -1;
    }
    const i64 blocks = std::min<i64>((n / 2 + 255) / 256, 8192);
    hipLaunchKernelGGL(k_copy, dim3((unsigned)std::max<i64>(blocks, 1)), dim3(256), 0,
                       (hipStream_t)stream, dst, src, n);
    HIP_CHECK(hipGetLastError());
    return 0;
}

int chp_logdiag(chp_plan* P, const double* lblkval, double* out, void* stream) {
    if (!P || !lblkval || !out) { chp_set_error("null argument"); return -1; }
    if (P->n == 0) return 0;
    hipLaunchKernelGGL(k_logdiag, dim3((unsigned)((P->n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, P->d_colidx, (int)P->n, P->d_fi, P->d_colfront,
                       lblkval, out);
    HIP_CHECK(hipGetLastError());
    return 0;
}

int chp_frontal_add_update(double* F, i64 N, const i64* relidx, const double* U,
                           const i64* strides, double alpha, void* stream) {
    if (N <= 0) return 0;
    hipLaunchKernelGGL(k_frontal_add, dim3((unsigned)N), dim3(256), 0, (hipStream_t)stream, F,
                       N, relidx, U, strides[0], strides[1], strides[2], strides[3], alpha);
    HIP_CHECK(hipGetLastError());
    return 0;
}

int chp_frontal_get_update(const double* F, i64 N, const i64* relidx, double* U,
                           const i64* strides, void* stream) {
    if (N <= 0) return 0;
    hipLaunchKernelGGL(k_frontal_get, dim3((unsigned)N), dim3(256), 0, (hipStream_t)stream, F,
                       N, relidx, U, strides[0], strides[1], strides[2], strides[3]);
    HIP_CHECK(hipGetLastError());
    return 0;
}

}  // extern "C"
