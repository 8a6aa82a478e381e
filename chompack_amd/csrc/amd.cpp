// Approximate minimum degree ordering (host, integer, deterministic).
//
// CHOMPACK orders with cvxopt.amd.order (symbolic.py:540-546), a wrapper of
// SuiteSparse AMD (Amestoy, Davis, Duff: "An approximate minimum degree
// ordering algorithm", SIAM J. Matrix Anal. Appl. 17(4), 1996; "Algorithm 837:
// AMD", ACM TOMS 30(3), 2004).  cvxopt is not available here, so this file
// restates the published algorithm with the library's default controls
// (dense threshold alpha = 10, aggressive absorption on) and its documented
// tie-breaking: degree lists are LIFO, the pivot is the head of the lowest
// non-empty degree list, hash buckets are LIFO, the final order is a
// postorder of the assembly tree with the largest child (by front size) last,
// and variables merged into a supervariable are ordered just before its
// principal variable, in index order.
//
// Quotient graph storage: one integer pool `iw`; object i owns
// iw[pe[i] .. pe[i]+len[i]-1], whose first elen[i] entries are elements and
// the rest variables.  A new element is assembled in a scratch vector and then
// appended to the pool, compacting the pool first when it is full.
#include <cstdint>
#include <cmath>
#include <vector>
#include <algorithm>

typedef int64_t i64;

namespace {

const i64 EMPTY = -1;
inline i64 flip(i64 x) { return -x - 2; }

struct Amd {
    i64 n;
    std::vector<i64> pe, len, elen, nv, degree, head, next, last, hhead, w, iw;
    i64 pfree = 0;

    // Compact the pool: copy every live list (pe >= 0, except `skip`) to the
    // front in order of position.  List contents and order are unchanged.
    void compact(i64 skip) {
        std::vector<std::pair<i64, i64>> objs;
        for (i64 j = 0; j < n; ++j)
            if (j != skip && pe[j] >= 0) objs.push_back({pe[j], j});
        std::sort(objs.begin(), objs.end());
        i64 dst = 0;
        for (auto& o : objs) {
            i64 j = o.second, src = pe[j];
            pe[j] = dst;
            for (i64 t = 0; t < len[j]; ++t) iw[dst + t] = iw[src + t];
            dst += len[j];
        }
        pfree = dst;
    }

    void remove_from_degree_list(i64 i) {
        i64 ilast = last[i], inext = next[i];
        if (inext != EMPTY) last[inext] = ilast;
        if (ilast != EMPTY) next[ilast] = inext;
        else head[degree[i]] = inext;
    }
};

}  // namespace

extern "C" {

// Order a symmetric matrix given by its full pattern in compressed-column form
// (cp[n+1], ri[nnz]; diagonal entries and duplicates are ignored).
// Writes the permutation to perm[n] (perm[k] = index of the k-th pivot).
// Returns 0 on success, -1 on invalid input.
int chp_amd_order(i64 n, const i64* cp, const i64* ri, i64* perm) {
    if (n < 0) return -1;
    if (n == 0) return 0;
    Amd a;
    a.n = n;
    // adjacency lists (sorted ascending, no diagonal, no duplicates)
    std::vector<i64> deg0(n, 0);
    std::vector<std::vector<i64>> adj(n);
    for (i64 j = 0; j < n; ++j)
        for (i64 p = cp[j]; p < cp[j + 1]; ++p) {
            i64 i = ri[p];
            if (i < 0 || i >= n) return -1;
            if (i == j) continue;
            adj[i].push_back(j);
            adj[j].push_back(i);
        }
    i64 nz = 0;
    for (i64 j = 0; j < n; ++j) {
        auto& v = adj[j];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        nz += (i64)v.size();
    }
    i64 iwlen = nz + nz / 5 + 2 * n + 16;
    a.pe.assign(n, 0); a.len.assign(n, 0); a.elen.assign(n, 0); a.nv.assign(n, 1);
    a.degree.assign(n, 0); a.head.assign(n, EMPTY); a.next.assign(n, EMPTY);
    a.last.assign(n, EMPTY); a.hhead.assign(n, EMPTY); a.w.assign(n, 1);
    a.iw.assign(iwlen, 0);
    i64 pos = 0;
    for (i64 j = 0; j < n; ++j) {
        a.pe[j] = pos;
        a.len[j] = (i64)adj[j].size();
        for (i64 v : adj[j]) a.iw[pos++] = v;
        a.degree[j] = a.len[j];
    }
    a.pfree = pos;
    adj.clear(); adj.shrink_to_fit();

    auto& pe = a.pe; auto& len = a.len; auto& elen = a.elen; auto& nv = a.nv;
    auto& degree = a.degree; auto& head = a.head; auto& next = a.next;
    auto& last = a.last; auto& hhead = a.hhead; auto& w = a.w; auto& iw = a.iw;

    double dalpha = 10.0;
    i64 dense = (i64)(dalpha * std::sqrt((double)n));
    dense = std::max<i64>(16, dense);
    dense = std::min<i64>(n, dense);

    i64 wflg = 2, mindeg = 0, nel = 0, lemax = 0;
    for (i64 i = 0; i < n; ++i) {
        i64 d = degree[i];
        if (d == 0) {
            elen[i] = flip(1); ++nel; pe[i] = EMPTY; w[i] = 0;
        } else if (d > dense) {
            nv[i] = 0; elen[i] = EMPTY; ++nel; pe[i] = EMPTY;
        } else {
            i64 inext = head[d];
            if (inext != EMPTY) last[inext] = i;
            next[i] = inext;
            head[d] = i;
        }
    }

    std::vector<i64> lme;
    while (nel < n) {
        // pivot of minimum degree
        i64 d = mindeg, me = EMPTY;
        for (; d < n; ++d) { me = head[d]; if (me != EMPTY) break; }
        if (me == EMPTY) return -1;
        mindeg = d;
        i64 inext = next[me];
        if (inext != EMPTY) last[inext] = EMPTY;
        head[d] = inext;
        i64 elenme = elen[me], nvpiv = nv[me];
        nel += nvpiv;

        // new element Lme
        nv[me] = -nvpiv;
        i64 degme = 0, pme1, pme2;
        if (elenme == 0) {
            pme1 = pe[me];
            pme2 = pme1 - 1;
            for (i64 p = pme1; p < pme1 + len[me]; ++p) {
                i64 i = iw[p], nvi = nv[i];
                if (nvi > 0) {
                    degme += nvi; nv[i] = -nvi; iw[++pme2] = i;
                    a.remove_from_degree_list(i);
                }
            }
        } else {
            lme.clear();
            i64 p = pe[me];
            i64 slenme = len[me] - elenme;
            for (i64 k1 = 1; k1 <= elenme + 1; ++k1) {
                i64 e, pj, ln;
                if (k1 > elenme) { e = me; pj = p; ln = slenme; }
                else { e = iw[p++]; pj = pe[e]; ln = len[e]; }
                for (i64 k2 = 0; k2 < ln; ++k2) {
                    i64 i = iw[pj++], nvi = nv[i];
                    if (nvi > 0) {
                        degme += nvi; nv[i] = -nvi; lme.push_back(i);
                        a.remove_from_degree_list(i);
                    }
                }
                if (e != me) { pe[e] = flip(me); w[e] = 0; }
            }
            if (a.pfree + (i64)lme.size() > iwlen) {
                a.compact(me);
                if (a.pfree + (i64)lme.size() > iwlen) {
                    iwlen = a.pfree + (i64)lme.size() + n;
                    iw.resize(iwlen);
                }
            }
            pme1 = a.pfree;
            for (i64 v : lme) iw[a.pfree++] = v;
            pme2 = a.pfree - 1;
        }
        degree[me] = degme;
        pe[me] = pme1;
        len[me] = pme2 - pme1 + 1;
        elen[me] = flip(nvpiv + degme);

        // |Le \ Lme| for every element adjacent to a variable of Lme
        for (i64 pme = pme1; pme <= pme2; ++pme) {
            i64 i = iw[pme], eln = elen[i];
            if (eln > 0) {
                i64 nvi = -nv[i], wnvi = wflg - nvi;
                for (i64 p = pe[i]; p < pe[i] + eln; ++p) {
                    i64 e = iw[p], we = w[e];
                    if (we >= wflg) we -= nvi;
                    else if (we != 0) we = degree[e] + wnvi;
                    w[e] = we;
                }
            }
        }

        // degree update, aggressive element absorption, hashing
        for (i64 pme = pme1; pme <= pme2; ++pme) {
            i64 i = iw[pme];
            i64 p1 = pe[i], p2 = p1 + elen[i] - 1, pn = p1;
            i64 hash = 0, dg = 0;
            for (i64 p = p1; p <= p2; ++p) {
                i64 e = iw[p], we = w[e];
                if (we != 0) {
                    i64 dext = we - wflg;
                    if (dext > 0) { dg += dext; iw[pn++] = e; hash += e; }
                    else { pe[e] = flip(me); w[e] = 0; }
                }
            }
            elen[i] = pn - p1 + 1;
            i64 p3 = pn, p4 = p1 + len[i];
            for (i64 p = p2 + 1; p < p4; ++p) {
                i64 j = iw[p], nvj = nv[j];
                if (nvj > 0) { dg += nvj; iw[pn++] = j; hash += j; }
            }
            if (elen[i] == 1 && p3 == pn) {
                // mass elimination
                pe[i] = flip(me);
                i64 nvi = -nv[i];
                degme -= nvi; nvpiv += nvi; nel += nvi;
                nv[i] = 0; elen[i] = EMPTY;
            } else {
                degree[i] = std::min(degree[i], dg);
                iw[pn] = iw[p3];
                iw[p3] = iw[p1];
                iw[p1] = me;
                len[i] = pn - p1 + 1;
                hash %= n;
                next[i] = hhead[hash];
                hhead[hash] = i;
                last[i] = hash;
            }
        }
        degree[me] = degme;
        lemax = std::max(lemax, degme);
        wflg += lemax;

        // supervariable detection
        for (i64 pme = pme1; pme <= pme2; ++pme) {
            i64 i = iw[pme];
            if (nv[i] >= 0) continue;
            i64 hash = last[i];
            i = hhead[hash];
            hhead[hash] = EMPTY;
            while (i != EMPTY && next[i] != EMPTY) {
                i64 ln = len[i], eln = elen[i];
                for (i64 p = pe[i] + 1; p < pe[i] + ln; ++p) w[iw[p]] = wflg;
                i64 jlast = i, j = next[i];
                while (j != EMPTY) {
                    bool ok = (len[j] == ln) && (elen[j] == eln);
                    for (i64 p = pe[j] + 1; ok && p < pe[j] + ln; ++p)
                        if (w[iw[p]] != wflg) ok = false;
                    if (ok) {
                        pe[j] = flip(i);
                        nv[i] += nv[j];
                        nv[j] = 0;
                        elen[j] = EMPTY;
                        j = next[j];
                        next[jlast] = j;
                    } else {
                        jlast = j;
                        j = next[j];
                    }
                }
                ++wflg;
                i = next[i];
            }
        }

        // restore degree lists, drop non-principal variables from Lme
        i64 p = pme1, nleft = n - nel;
        for (i64 pme = pme1; pme <= pme2; ++pme) {
            i64 i = iw[pme], nvi = -nv[i];
            if (nvi > 0) {
                nv[i] = nvi;
                i64 dg = std::min(degree[i] + degme - nvi, nleft - nvi);
                i64 hn = head[dg];
                if (hn != EMPTY) last[hn] = i;
                next[i] = hn;
                last[i] = EMPTY;
                head[dg] = i;
                mindeg = std::min(mindeg, dg);
                degree[i] = dg;
                iw[p++] = i;
            }
        }
        nv[me] = nvpiv;
        len[me] = p - pme1;
        if (len[me] == 0) { pe[me] = EMPTY; w[me] = 0; }
        if (elenme != 0) a.pfree = p;
    }

    // assembly tree: parent of j is flip(pe[j]); front size is flip(elen[e])
    for (i64 i = 0; i < n; ++i) pe[i] = flip(pe[i]);
    for (i64 i = 0; i < n; ++i) elen[i] = flip(elen[i]);
    for (i64 i = 0; i < n; ++i) {
        if (nv[i] != 0) continue;
        i64 j = pe[i];
        if (j == EMPTY) continue;  // dense variable
        while (nv[j] == 0) j = pe[j];
        i64 e = j;
        j = i;
        while (nv[j] == 0) { i64 jn = pe[j]; pe[j] = e; j = jn; }
    }

    // postorder the assembly tree, biggest child (front size) last
    std::vector<i64> child(n, EMPTY), sibling(n, EMPTY), order(n, EMPTY), stack(n);
    for (i64 j = n - 1; j >= 0; --j) {
        if (nv[j] > 0 && pe[j] != EMPTY) {
            sibling[j] = child[pe[j]];
            child[pe[j]] = j;
        }
    }
    for (i64 i = 0; i < n; ++i) {
        if (nv[i] <= 0 || child[i] == EMPTY) continue;
        i64 fprev = EMPTY, maxfr = EMPTY, bigfprev = EMPTY, bigf = EMPTY;
        for (i64 f = child[i]; f != EMPTY; f = sibling[f]) {
            if (elen[f] >= maxfr) { maxfr = elen[f]; bigfprev = fprev; bigf = f; }
            fprev = f;
        }
        i64 fnext = sibling[bigf];
        if (fnext != EMPTY) {
            if (bigfprev == EMPTY) child[i] = fnext;
            else sibling[bigfprev] = fnext;
            sibling[bigf] = EMPTY;
            sibling[fprev] = bigf;
        }
    }
    i64 k = 0;
    for (i64 r = 0; r < n; ++r) {
        if (pe[r] != EMPTY || nv[r] <= 0) continue;
        i64 top = 0;
        stack[0] = r;
        while (top >= 0) {
            i64 i = stack[top];
            if (child[i] != EMPTY) {
                i64 cnt = 0;
                for (i64 f = child[i]; f != EMPTY; f = sibling[f]) ++cnt;
                i64 h = top + cnt;
                top = h;
                for (i64 f = child[i]; f != EMPTY; f = sibling[f]) stack[h--] = f;
                child[i] = EMPTY;
            } else {
                --top;
                order[i] = k++;
            }
        }
    }

    // positions: each element's block, merged variables first, principal last
    std::vector<i64> byorder(n, EMPTY), position(n, EMPTY);
    for (i64 e = 0; e < n; ++e)
        if (order[e] != EMPTY) byorder[order[e]] = e;
    i64 cnt = 0;
    for (i64 t = 0; t < n; ++t) {
        i64 e = byorder[t];
        if (e == EMPTY) break;
        position[e] = cnt;
        cnt += nv[e];
    }
    for (i64 i = 0; i < n; ++i) {
        if (nv[i] != 0) continue;
        i64 e = pe[i];
        if (e != EMPTY) { position[i] = position[e]; position[e]++; }
        else position[i] = cnt++;
    }
    for (i64 i = 0; i < n; ++i) {
        i64 q = position[i];
        if (q < 0 || q >= n) return -1;
        perm[q] = i;
    }
    return 0;
}

}  // extern "C"
