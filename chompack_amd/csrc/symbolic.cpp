// Host-side symbolic factorization kernels (integer, bit-exact).
//
// These routines restate the integer algorithms of CHOMPACK's symbolic layer
// (src/python/symbolic.py) in C++ so that analysis of n ~ 1e5..1e6 problems
// takes milliseconds instead of minutes.  Every routine produces exactly the
// arrays the reference produces for the same input pattern; the tie-breaking
// rules (child visiting order, "first child in postorder claims the parent",
// index-ordered supernode numbering) are kept on purpose because every
// downstream array (snpost, relidx, blkptr) depends on them.
//
// Conventions: all index arrays are int64 (the reference's int_t is
// Py_ssize_t); a tree root r has parent[r] == r.
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

typedef int64_t i64;

extern "C" {

// Elimination tree of a symmetric pattern given in compressed-column form
// (full pattern or upper part; only entries i < k of column k are used).
// Liu's algorithm with path compression through the ancestor array w.
// Follows symbolic.py:49-72.
void chp_etree(i64 n, const i64* cp, const i64* ri, i64* parent) {
    std::vector<i64> w(n);
    for (i64 k = 0; k < n; ++k) {
        parent[k] = k;
        w[k] = -1;
        for (i64 p = cp[k]; p < cp[k + 1]; ++p) {
            i64 i = ri[p];
            while (i != -1 && i < k) {
                i64 inext = w[i];
                w[i] = k;
                if (inext == -1) parent[i] = k;
                i = inext;
            }
        }
    }
}

// Postorder of a forest (roots r have parent[r] == r).  Children are visited
// in increasing index order and roots in increasing index order.
// Follows symbolic.py:6-47.
void chp_post_order(i64 n, const i64* parent, i64* post) {
    std::vector<i64> head(n, -1), next(n, 0), stack(n > 0 ? n : 1);
    for (i64 j = n - 1; j >= 0; --j) {
        if (parent[j] == j) continue;
        next[j] = head[parent[j]];
        head[parent[j]] = j;
    }
    i64 k = 0;
    for (i64 j = 0; j < n; ++j) {
        if (parent[j] != j) continue;
        i64 top = 0;
        stack[0] = j;
        while (top >= 0) {
            i64 p = stack[top];
            i64 i = head[p];
            if (i == -1) {
                --top;
                post[k++] = p;
            } else {
                head[p] = next[i];
                stack[++top] = i;
            }
        }
    }
}

// Column counts of the Cholesky factor (Gilbert-Ng-Peyton skeleton/leaf
// algorithm).  cp/ri is the full symmetric pattern; entries i > j of column j
// are used.  Follows symbolic.py:74-131.
void chp_counts(i64 n, const i64* cp, const i64* ri, const i64* parent,
                const i64* post, i64* colcount) {
    std::vector<i64> ancestor(n), maxfirst(n, -1), prevleaf(n, -1), first(n, -1);
    for (i64 i = 0; i < n; ++i) ancestor[i] = i;
    for (i64 k = 0; k < n; ++k) {
        i64 j = post[k];
        colcount[j] = (first[j] == -1) ? 1 : 0;
        while (j != -1 && first[j] == -1) {
            first[j] = k;
            j = parent[j];
        }
    }
    for (i64 k = 0; k < n; ++k) {
        i64 j = post[k];
        if (parent[j] != j) colcount[parent[j]] -= 1;
        for (i64 p = cp[j]; p < cp[j + 1]; ++p) {
            i64 i = ri[p];
            if (i <= j) continue;
            // is j a leaf of the i-th row subtree?
            if (first[j] <= maxfirst[i]) continue;
            maxfirst[i] = first[j];
            i64 jprev = prevleaf[i];
            prevleaf[i] = j;
            colcount[j] += 1;
            if (jprev == -1) continue;  // first leaf: q = i, no decrement
            i64 q = jprev;
            while (q != ancestor[q]) q = ancestor[q];
            i64 s = jprev;
            while (s != q) {
                i64 sp = ancestor[s];
                ancestor[s] = q;
                s = sp;
            }
            colcount[q] -= 1;
        }
        if (parent[j] != j) ancestor[j] = parent[j];
    }
    for (i64 j = 0; j < n; ++j)
        if (parent[j] != j) colcount[parent[j]] += colcount[j];
}

// Maximal (fundamental, Pothen-Sun) supernodes and the supernodal tree.
// A node's parent joins the node's supernode iff colcount[j]-1 ==
// colcount[parent] and the parent is not yet claimed; the first child in
// postorder wins.  Supernodes are numbered by their representative (lowest)
// vertex; members are listed in increasing order.
// Outputs: snode[n], snptr[nsn+1], snpar[nsn]; returns nsn.
// Follows symbolic.py:133-244.
i64 chp_supernodes(i64 n, const i64* parent, const i64* post, const i64* colcount,
                   i64* snode, i64* snptr, i64* snpar) {
    std::vector<i64> flag(n, -1), spar(n, -1);
    // children lists (order is irrelevant for the result, see the reference loop)
    std::vector<i64> chead(n, -1), cnext(n, -1);
    for (i64 j = n - 1; j >= 0; --j) {
        if (parent[j] == j) continue;
        cnext[j] = chead[parent[j]];
        chead[parent[j]] = j;
    }
    for (i64 idx = 0; idx < n; ++idx) {
        i64 j = post[idx];
        i64 pj = parent[j];
        if (pj != j) {
            if (colcount[j] - 1 == colcount[pj] && flag[pj] == -1) {
                if (flag[j] < 0) {
                    flag[pj] = j;
                    flag[j] -= 1;
                } else {
                    flag[pj] = flag[j];
                    flag[flag[j]] -= 1;
                }
            }
        } else {
            if (flag[j] < 0) spar[j] = j;
            else spar[flag[j]] = flag[j];
        }
        i64 k = flag[j] < 0 ? j : flag[j];
        for (i64 i = chead[j]; i != -1; i = cnext[i]) {
            i64 l = flag[i] < 0 ? i : flag[i];
            if (l != k) spar[l] = k;
        }
    }
    // number supernodes by representative index
    std::vector<i64> sn(n, -1);
    i64 nsn = 0;
    for (i64 i = 0; i < n; ++i)
        if (flag[i] < 0) sn[i] = nsn++;
    // sizes
    std::vector<i64> cnt(nsn, 0);
    for (i64 i = 0; i < n; ++i) cnt[sn[flag[i] < 0 ? i : flag[i]]]++;
    snptr[0] = 0;
    for (i64 k = 0; k < nsn; ++k) snptr[k + 1] = snptr[k] + cnt[k];
    std::vector<i64> fill(snptr, snptr + nsn);
    for (i64 i = 0; i < n; ++i) {
        i64 k = sn[flag[i] < 0 ? i : flag[i]];
        snode[fill[k]++] = i;
    }
    for (i64 i = 0; i < n; ++i)
        if (flag[i] < 0) snpar[sn[i]] = sn[spar[i]];
    return nsn;
}

// Clique row sets of a postordered, contiguously numbered supernodal
// factorization (snode == identity, snpost == identity).  cpl/ril is the
// lower-triangular pattern of the permuted matrix with sorted row indices.
// Clique k = sorted union of the patterns of its columns and of its children's
// separators; returns 0 on success, -1-k if clique k's size disagrees with
// colcount of its representative (sncolptr must already hold the sizes as a
// prefix sum).  Follows symbolic.py:328-384 (embed with lmerge).
i64 chp_embed(i64 nsn, const i64* cpl, const i64* ril, const i64* snptr,
              const i64* snpar, const i64* sncolptr, i64* snrowidx) {
    std::vector<i64> cnnz(nsn, 0), tmp;
    auto merge_into = [&](i64* dst, i64 nd, const i64* src, i64 ns) -> i64 {
        tmp.resize(nd + ns);
        i64 a = 0, b = 0, k = 0;
        while (a < nd && b < ns) {
            if (dst[a] < src[b]) tmp[k++] = dst[a++];
            else if (dst[a] > src[b]) tmp[k++] = src[b++];
            else { tmp[k++] = dst[a++]; ++b; }
        }
        while (a < nd) tmp[k++] = dst[a++];
        while (b < ns) tmp[k++] = src[b++];
        return k;
    };
    for (i64 k = 0; k < nsn; ++k) {
        i64 cap = sncolptr[k + 1] - sncolptr[k];
        i64* dst = snrowidx + sncolptr[k];
        i64 c0 = snptr[k];
        i64 nk = cpl[c0 + 1] - cpl[c0];
        if (nk > cap) return -1 - k;
        std::memcpy(dst, ril + cpl[c0], sizeof(i64) * nk);
        cnnz[k] = nk;
        for (i64 c = c0 + 1; c < snptr[k + 1]; ++c) {
            i64 m = merge_into(dst, cnnz[k], ril + cpl[c], cpl[c + 1] - cpl[c]);
            if (m > cap) return -1 - k;
            std::memcpy(dst, tmp.data(), sizeof(i64) * m);
            cnnz[k] = m;
        }
    }
    for (i64 k = 0; k < nsn; ++k) {
        i64 p = snpar[k];
        if (p == k) continue;
        i64 nn = snptr[k + 1] - snptr[k];
        i64* dst = snrowidx + sncolptr[p];
        i64 m = merge_into(dst, cnnz[p], snrowidx + sncolptr[k] + nn, cnnz[k] - nn);
        if (m > sncolptr[p + 1] - sncolptr[p]) return -1 - p;
        std::memcpy(dst, tmp.data(), sizeof(i64) * m);
        cnnz[p] = m;
    }
    for (i64 k = 0; k < nsn; ++k)
        if (cnnz[k] != sncolptr[k + 1] - sncolptr[k]) return -1 - k;
    return 0;
}

// Merge of two ascending lists into out[0..nl+nr) (cbase.c:64-101): the
// smaller head is taken; equal heads are written once and both advance, so
// a value repeated inside one list stays repeated.  Returns the merged
// length.
i64 chp_lmerge(const i64* left, const i64* right, i64 nl, i64 nr, i64* out) {
    i64 a = 0, b = 0, k = 0;
    while (a < nl && b < nr) {
        if (left[a] < right[b]) out[k++] = left[a++];
        else if (left[a] > right[b]) out[k++] = right[b++];
        else { out[k++] = left[a++]; ++b; }
    }
    while (a < nl) out[k++] = left[a++];
    while (b < nr) out[k++] = right[b++];
    return k;
}

// Relative indices: for each non-root supernode k with parent p, the
// positions of k's separator rows inside p's clique.  relptr[nsn+1] must be
// prefilled with the prefix sums of the separator sizes.
// Returns 0, or -1-k if a separator row of k is missing from its parent.
// Follows symbolic.py:386-411.
i64 chp_relative_idx(i64 nsn, const i64* snptr, const i64* snpar,
                     const i64* sncolptr, const i64* snrowidx,
                     const i64* relptr, i64* relidx) {
    for (i64 k = 0; k < nsn; ++k) {
        i64 p = snpar[k];
        i64 nn = snptr[k + 1] - snptr[k];
        const i64* a = snrowidx + sncolptr[k] + nn;
        i64 na = relptr[k + 1] - relptr[k];
        if (p == k) {
            if (na != 0) return -1 - k;
            continue;
        }
        const i64* b = snrowidx + sncolptr[p];
        i64 nb = sncolptr[p + 1] - sncolptr[p];
        i64 pos = 0;
        for (i64 t = 0; t < na; ++t) {
            while (pos < nb && b[pos] != a[t]) ++pos;
            if (pos == nb) return -1 - k;
            relidx[relptr[k] + t] = pos;
            ++pos;
        }
    }
    return 0;
}

// Maximum cardinality search (mcs.py:4-57): vertices are numbered from
// n - 1 down to 0, each time taking the unnumbered vertex with the most
// numbered neighbours.  Ties follow the reference exactly: the candidates of
// each weight form a stack (bucket lists S[w]), a vertex is pushed again each
// time its weight grows (stale entries are skipped when popped), and the
// initial stack S[0] holds 0 .. n-1 without ve, then ve on top.  cp/ri: the
// full symmetric pattern in compressed-column form, rows ascending in each
// column (the CCS of symmetrize(A)).  Writes the ordering p[0..n-1].
void chp_maxcardsearch(i64 n, const i64* cp, const i64* ri, i64 ve, i64* p) {
    std::vector<i64> w(n, 0);
    std::vector<std::vector<i64>> S(std::max<i64>(n, 1));
    for (i64 i = 0; i < n; ++i)
        if (i != ve) S[0].push_back(i);
    S[0].push_back(ve);
    i64 max_w = 0;
    for (i64 i = n - 1; i >= 0; --i) {
        i64 v = -1;
        for (;;) {
            if (!S[max_w].empty()) {
                v = S[max_w].back();
                S[max_w].pop_back();
                if (w[v] >= 0) break;
            } else {
                --max_w;
            }
        }
        p[i] = v;
        w[v] = -1;
        for (i64 q = cp[v]; q < cp[v + 1]; ++q) {
            const i64 r = ri[q];
            if (w[r] >= 0) {
                ++w[r];
                S[w[r]].push_back(r);
                max_w = std::max(max_w, w[r]);
            }
        }
    }
}

// Maximal chordal subgraph (maxchord.py:5-83, Dearing, Shier and Warner):
// the maximum cardinality search in which an unnumbered neighbour u of the
// vertex v just numbered gains v (and the edge {u, v}) only when u's set of
// chordal numbered neighbours C[u] is a subset of C[v].  Outputs the kept
// entries of the lower triangle in the reference's order (for each column j
// ascending, the rows in the order they were kept; a diagonal entry is kept
// when met), their positions in vals (so the caller copies the values), and
// the ordering p.  Returns the number of kept entries (at most cp[n]).
i64 chp_maxchord(i64 n, const i64* cp, const i64* ri, i64 ve, i64* p, i64* rows, i64* cols,
                 i64* src) {
    std::vector<i64> w(n, 0);
    std::vector<std::vector<i64>> S(std::max<i64>(n, 1));
    for (i64 i = 0; i < n; ++i)
        if (i != ve) S[0].push_back(i);
    S[0].push_back(ve);
    std::vector<std::vector<i64>> C(n);       // chordal numbered neighbours
    std::vector<std::vector<i64>> E(n), Q(n);  // kept rows and value positions per column
    std::vector<i64> mark(n, -1);
    i64 max_w = 0;
    for (i64 i = n - 1; i >= 0; --i) {
        i64 v = -1;
        for (;;) {
            if (!S[max_w].empty()) {
                v = S[max_w].back();
                S[max_w].pop_back();
                if (w[v] >= 0) break;
            } else {
                --max_w;
            }
        }
        p[i] = v;
        w[v] = -1;
        for (i64 x : C[v]) mark[x] = v;  // C[v] as a marked set; fixed while v's neighbours are visited
        for (i64 q = cp[v]; q < cp[v + 1]; ++q) {
            const i64 u = ri[q];
            if (w[u] >= 0) {
                bool subset = true;
                for (i64 x : C[u])
                    if (mark[x] != v) { subset = false; break; }
                if (subset) {
                    C[u].push_back(v);
                    ++w[u];
                    S[w[u]].push_back(u);
                    max_w = std::max(max_w, w[u]);
                    E[std::min(u, v)].push_back(std::max(u, v));
                    Q[std::min(u, v)].push_back(q);
                }
            } else if (u == v) {
                E[u].push_back(u);
                Q[u].push_back(q);
            }
        }
    }
    i64 cnt = 0;
    for (i64 j = 0; j < n; ++j)
        for (size_t t = 0; t < E[j].size(); ++t) {
            rows[cnt] = E[j][t];
            cols[cnt] = j;
            src[cnt] = Q[j][t];
            ++cnt;
        }
    return cnt;
}

}  // extern "C"
