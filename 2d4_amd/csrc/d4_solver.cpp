// d4_solver.cpp -- host sparse LU (see d4_solver.h).
#include "d4_solver.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

namespace d4host {

std::vector<int> min_degree_order(int n, const int* Ap, const int* Ai)
{
    // elimination graph as bit-set rows; degree = live neighbours
    const int W = (n + 63) / 64;
    std::vector<uint64_t> adj((size_t)n * W, 0ull), alive(W, 0ull), nb(W);
    auto setb = [&](int r, int c) { adj[(size_t)r * W + (c >> 6)] |= 1ull << (c & 63); };
    for (int c = 0; c < n; ++c)
        for (int k = Ap[c]; k < Ap[c + 1]; ++k) {
            const int r = Ai[k];
            if (r == c) continue;
            setb(r, c);
            setb(c, r);
        }
    for (int v = 0; v < n; ++v) alive[v >> 6] |= 1ull << (v & 63);
    std::vector<int> deg(n);
    for (int v = 0; v < n; ++v) {
        int d = 0;
        for (int w = 0; w < W; ++w) d += __builtin_popcountll(adj[(size_t)v * W + w]);
        deg[v] = d;
    }
    std::vector<char> done(n, 0);
    std::vector<int> order;
    order.reserve(n);
    for (int step = 0; step < n; ++step) {
        int v = -1;
        for (int u = 0; u < n; ++u)
            if (!done[u] && (v < 0 || deg[u] < deg[v])) v = u;
        order.push_back(v);
        done[v] = 1;
        alive[v >> 6] &= ~(1ull << (v & 63));
        const uint64_t* av = &adj[(size_t)v * W];
        for (int w = 0; w < W; ++w) nb[w] = av[w] & alive[w];
        for (int w = 0; w < W; ++w) {
            uint64_t bits = nb[w];
            while (bits) {
                const int u = w * 64 + __builtin_ctzll(bits);
                bits &= bits - 1;
                uint64_t* au = &adj[(size_t)u * W];
                int d = 0;
                for (int x = 0; x < W; ++x) {
                    au[x] = (au[x] | nb[x]) & alive[x];
                    d += 0;
                }
                au[u >> 6] &= ~(1ull << (u & 63));
                for (int x = 0; x < W; ++x) d += __builtin_popcountll(au[x]);
                deg[u] = d;
            }
        }
    }
    return order;
}

namespace {

// the entries in which two CSC patterns of order n differ (rows ascending in each column)
int64_t pattern_distance(int n, const int* Ap, const int* Ai, const std::vector<int>& Bp, const std::vector<int>& Bi)
{
    int64_t d = 0;
    for (int c = 0; c < n; ++c) {
        int p = Ap[c], q = Bp[c];
        const int pe = Ap[c + 1], qe = Bp[c + 1];
        while (p < pe && q < qe) {
            if (Ai[p] == Bi[q]) {
                ++p;
                ++q;
            } else {
                ++d;
                if (Ai[p] < Bi[q]) ++p;
                else ++q;
            }
        }
        d += (pe - p) + (qe - q);
    }
    return d;
}

// breadth-first levels of s's component in the graph whose neighbours of v
// are the rows of column v: the number of levels, and the nodes of the last
// one; stamp[v] == tag marks the nodes reached
int bfs_depth(int s, const int* Ap, const int* Ai, std::vector<int>& stamp, int tag, std::vector<int>& queue,
              std::vector<int>& last_level)
{
    queue.clear();
    queue.push_back(s);
    stamp[s] = tag;
    size_t begin = 0;
    int depth = 0;
    while (begin < queue.size()) {
        const size_t end = queue.size();
        ++depth;
        for (size_t h = begin; h < end; ++h) {
            const int v = queue[h];
            for (int k = Ap[v]; k < Ap[v + 1]; ++k)
                if (stamp[Ai[k]] != tag) {
                    stamp[Ai[k]] = tag;
                    queue.push_back(Ai[k]);
                }
        }
        if (queue.size() == end) last_level.assign(queue.begin() + begin, queue.begin() + end);
        begin = end;
    }
    return depth;
}

}  // namespace

std::vector<int> rcm_order(int n, const int* Ap, const int* Ai)
{
    auto deg = [&](int v) { return Ap[v + 1] - Ap[v]; };
    std::vector<int> order, queue, last, last2, stamp(n, -1);
    order.reserve(n);
    std::vector<char> seen(n, 0);
    int tag = 0;
    for (int s0 = 0; s0 < n; ++s0) {
        if (seen[s0]) continue;
        // pseudo-peripheral start: move to a lowest-degree node of the last
        // level while that lengthens the level structure
        int v = s0;
        int depth = bfs_depth(v, Ap, Ai, stamp, tag++, queue, last);
        for (int it = 0; it < 16; ++it) {
            int u = last[0];
            for (int w : last)
                if (deg(w) < deg(u) || (deg(w) == deg(u) && w < u)) u = w;
            const int d2 = bfs_depth(u, Ap, Ai, stamp, tag++, queue, last2);
            if (d2 <= depth) break;
            v = u;
            depth = d2;
            last.swap(last2);
        }
        // Cuthill-McKee search from v, neighbours in row order
        size_t head = order.size();
        order.push_back(v);
        seen[v] = 1;
        while (head < order.size()) {
            const int x = order[head++];
            for (int k = Ap[x]; k < Ap[x + 1]; ++k)
                if (!seen[Ai[k]]) {
                    seen[Ai[k]] = 1;
                    order.push_back(Ai[k]);
                }
        }
    }
    std::reverse(order.begin(), order.end());
    return order;
}

// first, last and top from f(i), the lowest column of row i's entries (f(i) <= i, npad entries)
static BlockEnvelope envelope_from_rows(const std::vector<int>& f, int npad, int bs)
{
    BlockEnvelope e;
    const int nb = npad / bs;
    e.first.assign(nb, 0);
    e.last.assign(nb, 0);
    e.top.assign(nb, 0);
    for (int b = 0; b < nb; ++b) {
        int lo = f[b * bs];
        for (int i = b * bs + 1; i < (b + 1) * bs; ++i) lo = std::min(lo, f[i]);
        e.first[b] = lo / bs;
        e.last[b] = b;
    }
    for (int b = 0; b < nb; ++b) e.last[e.first[b]] = std::max(e.last[e.first[b]], b);
    for (int b = 1; b < nb; ++b) e.last[b] = std::max(e.last[b], e.last[b - 1]);
    for (int b = 0, p = 0; b < nb; ++b) {
        while (e.last[p] < b) ++p;
        e.top[b] = p;
    }
    for (int p = 0; p < nb; ++p) e.tiles += (long long)(e.last[p] - p) * (e.last[p] - p);
    return e;
}

BlockEnvelope block_envelope(int n, int npad, int bs, const int* Ap, const int* Ai, const int* inv, bool symmetric)
{
    std::vector<int> f(npad);
    for (int i = 0; i < npad; ++i) f[i] = i;
    if (symmetric) {
        // row i's entries are column i's (the pattern is symmetric): f(new c) is the lowest new index among
        // column c's rows, one gather pass and no scattered updates
        for (int c = 0; c < n; ++c) {
            const int nc = inv ? inv[c] : c;
            int lo = nc;
            if (inv)
                for (int k = Ap[c]; k < Ap[c + 1]; ++k) lo = std::min(lo, inv[Ai[k]]);
            else
                for (int k = Ap[c]; k < Ap[c + 1]; ++k) lo = std::min(lo, Ai[k]);
            f[nc] = lo;
        }
        return envelope_from_rows(f, npad, bs);
    }
    for (int c = 0; c < n; ++c) {
        const int nc = inv ? inv[c] : c;
        for (int k = Ap[c]; k < Ap[c + 1]; ++k) {
            const int nr = inv ? inv[Ai[k]] : Ai[k];
            const int lo = std::min(nr, nc), hi = std::max(nr, nc);
            f[hi] = std::min(f[hi], lo);
        }
    }
    return envelope_from_rows(f, npad, bs);
}

BlockEnvelope key_band_envelope(int n, int npad, int bs, const double* key, double reach)
{
    std::vector<int> f(npad);
    for (int i = 0; i < npad; ++i) f[i] = i;
    // row i can hold entries from the first row whose key is within reach below its own
    for (int i = 0; i < n; ++i)
        f[i] = std::min(i, (int)(std::lower_bound(key, key + n, key[i] - reach) - key));
    return envelope_from_rows(f, npad, bs);
}

bool band_order(int n, int npad, int bs, const int* Ap, const int* Ai, std::vector<int>& ord, std::vector<int>& inv,
                BlockEnvelope& env, bool symmetric)
{
    env = block_envelope(n, npad, bs, Ap, Ai, nullptr, symmetric);
    ord.clear();
    inv.clear();
    if (n < 2 * bs || env.tiles == 0) return false;
    std::vector<int> o = rcm_order(n, Ap, Ai);
    std::vector<int> iv(n);
    for (int i = 0; i < n; ++i) iv[o[i]] = i;
    BlockEnvelope e = block_envelope(n, npad, bs, Ap, Ai, iv.data(), symmetric);
    if (2 * e.tiles > env.tiles) return false;
    ord.swap(o);
    inv.swap(iv);
    env = std::move(e);
    return true;
}

int SparseLU::factor_dense(int n, const int* Ap, const int* Ai, const double* Ax)
{
    dense_ = true;
    D_.assign((size_t)n * n, 0.0);
    dpiv_.assign(n, 0);
    for (int c = 0; c < n; ++c)
        for (int k = Ap[c]; k < Ap[c + 1]; ++k) D_[(size_t)c * n + Ai[k]] = Ax[k];
    double* a = D_.data();
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = std::fabs(a[(size_t)k * n + k]);
        for (int r = k + 1; r < n; ++r) {
            const double v = std::fabs(a[(size_t)k * n + r]);
            if (v > best) {
                best = v;
                p = r;
            }
        }
        dpiv_[k] = p;
        if (p != k)
            for (int c = 0; c < n; ++c) std::swap(a[(size_t)c * n + k], a[(size_t)c * n + p]);
        const double d = a[(size_t)k * n + k];
        if (d == 0.0) {
            singular_ = true;
            continue;
        }
        const double inv = 1.0 / d;
        for (int r = k + 1; r < n; ++r) a[(size_t)k * n + r] *= inv;
        for (int c = k + 1; c < n; ++c) {
            double* col = a + (size_t)c * n;
            const double m = col[k];
            if (m == 0.0) continue;
            const double* lk = a + (size_t)k * n;
            for (int r = k + 1; r < n; ++r) col[r] -= lk[r] * m;
        }
    }
    return singular_ ? 1 : 0;
}

int SparseLU::factor(int n, const int* Ap, const int* Ai, const double* Ax)
{
    if (n < 0 || !Ap || (n > 0 && (!Ai || !Ax))) return -1;
    n_ = n;
    singular_ = false;
    dense_ = false;
    const int64_t nnz = Ap[n];
    if (n <= 96 || nnz > (int64_t)n * n / 4) return factor_dense(n, Ap, Ai, Ax);

    // The pattern of the previous factorization again (the rescaled stage-II
    // matrix, a retried step): replay its reach.  Otherwise the column order
    // (any column order is valid) is kept while the pattern differs in at most
    // 1 % of its entries from the one it was computed for -- a few pairs cross
    // the cutoff between Jacobians -- and computed afresh beyond that.
    const bool same = n == pat_n_ && (int64_t)pat_i_.size() == nnz && std::equal(Ap, Ap + n + 1, pat_p_.begin()) &&
                      std::equal(Ai, Ai + nnz, pat_i_.begin());
    if (!same) {
        if (!(n == ord_n_ && 100 * pattern_distance(n, Ap, Ai, ord_p_, ord_i_) <= nnz)) {
            q_ = min_degree_order(n, Ap, Ai);
            ++orderings_;
            ord_n_ = n;
            ord_p_.assign(Ap, Ap + n + 1);
            ord_i_.assign(Ai, Ai + nnz);
        }
        pat_n_ = n;
        pat_p_.assign(Ap, Ap + n + 1);
        pat_i_.assign(Ai, Ai + nnz);
        have_reach_ = false;
    } else if (have_reach_ && refactor(n, Ap, Ai, Ax)) {
        ++refactorizations_;
        return singular_ ? 1 : 0;
    }
    have_reach_ = false;
    reach_p_.assign(n + 1, 0);
    reach_i_.clear();
    pinv_.assign(n, -1);
    prow_.assign(n, -1);
    Lp_.assign(n + 1, 0);
    Up_.assign(n + 1, 0);
    Li_.clear();
    Lx_.clear();
    Ui_.clear();
    Ux_.clear();
    Li_.reserve((size_t)nnz * 4);
    Lx_.reserve((size_t)nnz * 4);
    Ui_.reserve((size_t)nnz * 4);
    Ux_.reserve((size_t)nnz * 4);

    std::vector<double> x(n, 0.0);
    std::vector<int> stack(n), pstack(n), xi(n);
    std::vector<char> mark(n, 0);
    bool structural = false;
    for (int k = 0; k < n; ++k) {
        const int col = q_[k];
        // reach of the column's pattern in the graph of L (pivot-indexed), topological order
        int top = n;
        for (int p = Ap[col]; p < Ap[col + 1]; ++p) {
            const int r0 = Ai[p];
            if (mark[r0]) continue;
            int head = 0;
            stack[0] = r0;
            while (head >= 0) {
                const int j = stack[head];
                const int jc = pinv_[j];  // L column of row j, if j is pivotal
                if (!mark[j]) {
                    mark[j] = 1;
                    pstack[head] = jc < 0 ? 0 : Lp_[jc];
                }
                bool pushed = false;
                if (jc >= 0) {
                    const int end = Lp_[jc + 1];
                    for (int t = pstack[head]; t < end; ++t) {
                        const int i = Li_[t];
                        if (mark[i]) continue;
                        pstack[head] = t + 1;
                        stack[++head] = i;
                        pushed = true;
                        break;
                    }
                }
                if (!pushed) {
                    head--;
                    xi[--top] = j;
                }
            }
        }
        reach_i_.insert(reach_i_.end(), xi.begin() + top, xi.end());
        reach_p_[k + 1] = (int)reach_i_.size();
        // x = L \ A(:, col) over the reach
        for (int p = Ap[col]; p < Ap[col + 1]; ++p) x[Ai[p]] = Ax[p];
        for (int t = top; t < n; ++t) {
            const int j = xi[t];
            const int jc = pinv_[j];
            if (jc < 0) continue;
            const double xj = x[j];
            for (int s = Lp_[jc]; s < Lp_[jc + 1]; ++s) x[Li_[s]] -= Lx_[s] * xj;
        }
        // pivot: diagonal preferred (row col), else largest
        int ipiv = -1;
        double amax = -1.0;
        for (int t = top; t < n; ++t) {
            const int i = xi[t];
            if (pinv_[i] < 0) {
                const double v = std::fabs(x[i]);
                if (v > amax) {
                    amax = v;
                    ipiv = i;
                }
            } else {
                Ui_.push_back(pinv_[i]);
                Ux_.push_back(x[i]);
            }
        }
        if (ipiv < 0) {  // structurally singular column
            singular_ = true;
            structural = true;
            for (int i = 0; i < n; ++i)
                if (pinv_[i] < 0) {
                    ipiv = i;
                    break;
                }
            amax = 0.0;
            x[ipiv] = 0.0;
        }
        if (pinv_[col] < 0 && mark[col] && std::fabs(x[col]) >= 0.1 * amax) ipiv = col;
        const double piv = x[ipiv];
        if (piv == 0.0) singular_ = true;
        Ui_.push_back(k);
        Ux_.push_back(piv);
        Up_[k + 1] = (int)Ui_.size();
        pinv_[ipiv] = k;
        prow_[k] = ipiv;
        const double inv = piv != 0.0 ? 1.0 / piv : 0.0;
        for (int t = top; t < n; ++t) {
            const int i = xi[t];
            if (pinv_[i] < 0) {
                Li_.push_back(i);
                Lx_.push_back(x[i] * inv);
            }
            x[i] = 0.0;
            mark[i] = 0;
        }
        if (!mark[ipiv]) x[ipiv] = 0.0;
        Lp_[k + 1] = (int)Li_.size();
    }
    // L row indices from original rows to pivot order (kept in original rows for refactor())
    Lrow_ = Li_;
    for (auto& i : Li_) i = pinv_[i];
    have_reach_ = !structural;
    return singular_ ? 1 : 0;
}

bool SparseLU::refactor(int n, const int* Ap, const int* Ai, const double* Ax)
{
    // The pattern, the column order and (checked column by column) the pivot
    // sequence are the last factorization's, so each column's reach and its
    // topological order are too: replay them, with the same operations in the
    // same order as factor() (bitwise its results).  false: a pivot choice
    // differs, and factor() starts over.
    singular_ = false;
    work_.assign(n, 0.0);
    double* x = work_.data();
    size_t lpos = 0, upos = 0;
    for (int k = 0; k < n; ++k) {
        const int col = q_[k];
        const int* R = reach_i_.data() + reach_p_[k];
        const int nr = reach_p_[k + 1] - reach_p_[k];
        for (int p = Ap[col]; p < Ap[col + 1]; ++p) x[Ai[p]] = Ax[p];
        for (int t = 0; t < nr; ++t) {
            const int jc = pinv_[R[t]];
            if (jc >= k) continue;  // not pivotal before column k
            const double xj = x[R[t]];
            for (int s = Lp_[jc]; s < Lp_[jc + 1]; ++s) x[Lrow_[s]] -= Lx_[s] * xj;
        }
        int ipiv = -1;
        double amax = -1.0;
        bool has_col = false;
        for (int t = 0; t < nr; ++t) {
            const int i = R[t];
            has_col |= i == col;
            if (pinv_[i] >= k) {
                const double v = std::fabs(x[i]);
                if (v > amax) {
                    amax = v;
                    ipiv = i;
                }
            } else {
                Ux_[upos++] = x[i];
            }
        }
        if (ipiv >= 0 && pinv_[col] >= k && has_col && std::fabs(x[col]) >= 0.1 * amax) ipiv = col;
        if (ipiv != prow_[k]) {
            for (int t = 0; t < nr; ++t) x[R[t]] = 0.0;
            for (int p = Ap[col]; p < Ap[col + 1]; ++p) x[Ai[p]] = 0.0;
            return false;
        }
        const double piv = x[ipiv];
        if (piv == 0.0) singular_ = true;
        Ux_[upos++] = piv;
        const double inv = piv != 0.0 ? 1.0 / piv : 0.0;
        for (int t = 0; t < nr; ++t) {
            const int i = R[t];
            if (pinv_[i] > k) Lx_[lpos++] = x[i] * inv;
            x[i] = 0.0;
        }
        x[ipiv] = 0.0;
    }
    return true;
}

void SparseLU::solve_const(const double* b, double* xout) const
{
    const int n = n_;
    if (dense_) {
        std::vector<double> c(b, b + n);
        for (int k = 0; k < n; ++k) std::swap(c[k], c[dpiv_[k]]);
        const double* a = D_.data();
        for (int k = 0; k < n; ++k) {
            const double v = c[k];
            if (v != 0.0)
                for (int r = k + 1; r < n; ++r) c[r] -= a[(size_t)k * n + r] * v;
        }
        for (int k = n - 1; k >= 0; --k) {
            c[k] /= a[(size_t)k * n + k];
            const double v = c[k];
            if (v != 0.0)
                for (int r = 0; r < k; ++r) c[r] -= a[(size_t)k * n + r] * v;
        }
        std::memcpy(xout, c.data(), sizeof(double) * (size_t)n);
        return;
    }
    std::vector<double> c(n);
    for (int i = 0; i < n; ++i) c[pinv_[i]] = b[i];
    for (int k = 0; k < n; ++k) {
        const double v = c[k];
        if (v != 0.0)
            for (int s = Lp_[k]; s < Lp_[k + 1]; ++s) c[Li_[s]] -= Lx_[s] * v;
    }
    for (int k = n - 1; k >= 0; --k) {
        const int last = Up_[k + 1] - 1;  // diagonal stored last
        c[k] /= Ux_[last];
        const double v = c[k];
        if (v != 0.0)
            for (int s = Up_[k]; s < last; ++s) c[Ui_[s]] -= Ux_[s] * v;
    }
    for (int k = 0; k < n; ++k) xout[q_[k]] = c[k];
}

}  // namespace d4host
