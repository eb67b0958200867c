// metriko (MI355X build) -- COLAMD column ordering, as Eigen 3.4's COLAMDOrdering computes it.
//
// reduce_to_linearly_independent (reference solver/matrix_ops.h:12-28) factorises C^T with
// Eigen::SparseQR<SprsD, COLAMDOrdering<int>>; the rows it keeps, and their order, follow the
// column order COLAMD gives C^T.  This is a restatement of the published algorithm (Davis,
// Gilbert, Larimore, Ng, "Algorithm 836: COLAMD, a column approximate minimum degree ordering
// algorithm", ACM TOMS 30(3), 2004) with the defaults of Eigen's port: a row is dense above
// 0.5 n_col entries, a column above 0.5 n_row, and rows whose set difference with the pivot
// row becomes empty are absorbed.  Steps: build the row form; order empty, then dense,
// columns last; score each column by the sum of its rows' external degrees; then repeatedly
// take the column of least approximate external degree, merge its rows into a new pivot row,
// update the degrees of that row's columns from the set differences of their rows, eliminate
// columns left with no other row, merge identical columns (hash on row-index sums) into
// super-columns, and finally order each absorbed column right after its principal column.
// Storage compaction (the published garbage collection) only removes dead entries that every
// scan skips, so the row store simply grows here.
#ifndef METRIKO_SOLVER_COLAMD_H
#define METRIKO_SOLVER_COLAMD_H
#include <algorithm>
#include <cstdint>
#include <vector>

namespace metriko {
namespace colamd {

// p[k] = the column placed k-th.  colptr / rowind: the pattern in compressed-column form, row
// indices sorted and unique in each column.
template <typename PtrVec, typename IdxVec>
std::vector<int> order(int n_row, int n_col, const PtrVec& colptr, const IdxVec& rowind) {
    constexpr int kEmpty = -1, kDeadPrincipal = -1, kDeadNonPrincipal = -2;
    const int64_t nnz = colptr[size_t(n_col)];
    // entries: column lists first, then row lists, then the pivot rows as they are formed
    std::vector<int64_t> A(static_cast<size_t>(2 * nnz));
    for (int64_t k = 0; k < nnz; k++) A[size_t(k)] = rowind[size_t(k)];
    // column records; the "shared" fields keep their overlapping meanings
    std::vector<int64_t> cstart(static_cast<size_t>(n_col));
    std::vector<int64_t> clen(static_cast<size_t>(n_col));
    std::vector<int64_t> thick_or_parent(static_cast<size_t>(n_col)), score_or_order(static_cast<size_t>(n_col));
    std::vector<int64_t> prev_or_hash(static_cast<size_t>(n_col)), next_link(static_cast<size_t>(n_col));
    // row records
    std::vector<int64_t> rstart(static_cast<size_t>(n_row)), rlen(static_cast<size_t>(n_row)), degree(static_cast<size_t>(n_row)), mark(static_cast<size_t>(n_row));

    for (int c = 0; c < n_col; c++) {
        cstart[c] = colptr[c];
        clen[c] = colptr[c + 1] - colptr[c];
        thick_or_parent[c] = 1;
        score_or_order[c] = 0;
        prev_or_hash[c] = kEmpty;
        next_link[c] = kEmpty;
    }
    std::fill(rlen.begin(), rlen.end(), 0);
    std::fill(mark.begin(), mark.end(), -1);
    for (int c = 0; c < n_col; c++)
        for (int64_t k = colptr[c]; k < colptr[c + 1]; k++) {
            const int64_t r = A[size_t(k)];
            if (mark[r] != c) rlen[r]++;
            else clen[c]--;
            mark[r] = c;
        }
    std::vector<int64_t> fill(static_cast<size_t>(n_row));
    for (int r = 0; r < n_row; r++) {
        rstart[r] = r == 0 ? nnz : rstart[r - 1] + rlen[r - 1];
        fill[r] = rstart[r];
    }
    for (int c = 0; c < n_col; c++)
        for (int64_t k = colptr[c]; k < colptr[c + 1]; k++) A[size_t(fill[A[size_t(k)]]++)] = c;
    for (int r = 0; r < n_row; r++) {
        mark[r] = 0;
        degree[r] = rlen[r];
    }
    auto col_alive = [&](int64_t c) { return cstart[c] >= 0; };
    auto row_alive = [&](int64_t r) { return mark[r] >= 0; };

    // ---- initial scoring --------------------------------------------------------------------
    const int64_t dense_row = std::max<int64_t>(0, std::min<int64_t>(int64_t(0.5 * n_col), n_col));
    const int64_t dense_col = std::max<int64_t>(0, std::min<int64_t>(int64_t(0.5 * n_row), n_row));
    int64_t max_deg = 0, n_col2 = n_col;
    for (int c = n_col - 1; c >= 0; c--)
        if (clen[c] == 0) { score_or_order[c] = --n_col2; cstart[c] = kDeadPrincipal; }
    for (int c = n_col - 1; c >= 0; c--) {
        if (!col_alive(c) || clen[c] <= dense_col) continue;
        score_or_order[c] = --n_col2;
        for (int64_t k = cstart[c]; k < cstart[c] + clen[c]; k++) degree[A[size_t(k)]]--;
        cstart[c] = kDeadPrincipal;
    }
    for (int r = 0; r < n_row; r++) {
        if (degree[r] > dense_row || degree[r] == 0) mark[r] = -1;
        else max_deg = std::max(max_deg, degree[r]);
    }
    for (int c = n_col - 1; c >= 0; c--) {
        if (!col_alive(c)) continue;
        int64_t score = 0, w = cstart[c];
        for (int64_t k = cstart[c]; k < cstart[c] + clen[c]; k++) {
            const int64_t r = A[size_t(k)];
            if (!row_alive(r)) continue;
            A[size_t(w++)] = r;
            score = std::min<int64_t>(score + degree[r] - 1, n_col);
        }
        if (w == cstart[c]) { score_or_order[c] = --n_col2; cstart[c] = kDeadPrincipal; }
        else { clen[c] = w - cstart[c]; score_or_order[c] = score; }
    }
    std::vector<int64_t> head(size_t(n_col) + 1, kEmpty);
    int64_t min_score = n_col;
    for (int c = n_col - 1; c >= 0; c--) {  // low indices at the front of each degree list
        if (!col_alive(c)) continue;
        const int64_t s = score_or_order[c], nx = head[s];
        prev_or_hash[c] = kEmpty;
        next_link[c] = nx;
        if (nx != kEmpty) prev_or_hash[nx] = c;
        head[s] = c;
        min_score = std::min(min_score, s);
    }

    // ---- ordering ----------------------------------------------------------------------------
    int64_t tag = 1;
    for (int r = 0; r < n_row; r++) if (row_alive(r)) mark[r] = 0;
    min_score = 0;
    for (int64_t k = 0; k < n_col2;) {
        while (head[min_score] == kEmpty && min_score < n_col) min_score++;
        const int64_t pc = head[min_score];
        {
            const int64_t nx = next_link[pc];
            head[min_score] = nx;
            if (nx != kEmpty) prev_or_hash[nx] = kEmpty;
        }
        score_or_order[pc] = k;
        const int64_t pthick = thick_or_parent[pc];
        k += pthick;

        // pivot row: union of the live rows of the pivot column
        const int64_t prs = int64_t(A.size());
        int64_t pdeg = 0;
        thick_or_parent[pc] = -pthick;
        for (int64_t q = cstart[pc]; q < cstart[pc] + clen[pc]; q++) {
            const int64_t row = A[size_t(q)];
            if (!row_alive(row)) continue;
            for (int64_t t = rstart[row]; t < rstart[row] + rlen[row]; t++) {
                const int64_t col = A[size_t(t)];
                const int64_t th = thick_or_parent[col];
                if (th > 0 && col_alive(col)) {
                    thick_or_parent[col] = -th;
                    A.push_back(col);
                    pdeg += th;
                }
            }
        }
        thick_or_parent[pc] = pthick;
        max_deg = std::max(max_deg, pdeg);
        for (int64_t q = cstart[pc]; q < cstart[pc] + clen[pc]; q++) mark[A[size_t(q)]] = -1;
        const int64_t plen = int64_t(A.size()) - prs;
        const int64_t prow = plen > 0 ? A[size_t(cstart[pc])] : kEmpty;

        // set differences of the rows met through the pivot row's columns
        for (int64_t t = prs; t < prs + plen; t++) {
            const int64_t col = A[size_t(t)];
            const int64_t th = -thick_or_parent[col];
            thick_or_parent[col] = th;
            const int64_t cs = score_or_order[col], pv = prev_or_hash[col], nx = next_link[col];
            if (pv == kEmpty) head[cs] = nx;
            else next_link[pv] = nx;
            if (nx != kEmpty) prev_or_hash[nx] = pv;
            for (int64_t q = cstart[col]; q < cstart[col] + clen[col]; q++) {
                const int64_t row = A[size_t(q)];
                const int64_t rm = mark[row];
                if (rm < 0) continue;
                int64_t sd = rm - tag;
                if (sd < 0) sd = degree[row];
                sd -= th;
                if (sd == 0) mark[row] = -1;  // aggressive absorption
                else mark[row] = sd + tag;
            }
        }
        // column scores from the set differences; mass elimination; hash buckets
        for (int64_t t = prs; t < prs + plen; t++) {
            const int64_t col = A[size_t(t)];
            uint64_t hash = 0;
            int64_t cur = 0, w = cstart[col];
            for (int64_t q = cstart[col]; q < cstart[col] + clen[col]; q++) {
                const int64_t row = A[size_t(q)];
                const int64_t rm = mark[row];
                if (rm < 0) continue;
                A[size_t(w++)] = row;
                hash += uint64_t(row);
                cur = std::min<int64_t>(cur + rm - tag, n_col);
            }
            clen[col] = w - cstart[col];
            if (clen[col] == 0) {
                cstart[col] = kDeadPrincipal;
                pdeg -= thick_or_parent[col];
                score_or_order[col] = k;
                k += thick_or_parent[col];
            } else {
                score_or_order[col] = cur;
                const int64_t h = int64_t(hash % uint64_t(n_col + 1));
                const int64_t hc = head[h];
                int64_t first;
                if (hc > kEmpty) { first = prev_or_hash[hc]; prev_or_hash[hc] = col; }
                else { first = -(hc + 2); head[h] = -(col + 2); }
                next_link[col] = first;
                prev_or_hash[col] = h;
            }
        }
        // super-columns: identical live row lists with equal scores merge into the first
        for (int64_t t = prs; t < prs + plen; t++) {
            const int64_t col = A[size_t(t)];
            if (!col_alive(col)) continue;
            const int64_t h = prev_or_hash[col];
            const int64_t hc = head[h];
            const int64_t first = hc > kEmpty ? prev_or_hash[hc] : -(hc + 2);
            for (int64_t sc = first; sc != kEmpty; sc = next_link[sc]) {
                const int64_t len = clen[sc];
                int64_t pv = sc;
                for (int64_t c = next_link[sc]; c != kEmpty; c = next_link[c]) {
                    if (clen[c] != len || score_or_order[c] != score_or_order[sc]) { pv = c; continue; }
                    if (!std::equal(A.begin() + cstart[sc], A.begin() + cstart[sc] + len, A.begin() + cstart[c])) {
                        pv = c;
                        continue;
                    }
                    thick_or_parent[sc] += thick_or_parent[c];
                    thick_or_parent[c] = sc;
                    cstart[c] = kDeadNonPrincipal;
                    score_or_order[c] = kEmpty;
                    next_link[pv] = next_link[c];
                }
            }
            if (hc > kEmpty) prev_or_hash[hc] = kEmpty;
            else head[h] = kEmpty;
        }
        cstart[pc] = kDeadPrincipal;
        tag += max_deg + 1;  // every mark below the new tag reads "not seen"
        // final scores; the pivot row joins its columns and they re-enter the degree lists
        int64_t w = prs;
        for (int64_t t = prs; t < prs + plen; t++) {
            const int64_t col = A[size_t(t)];
            if (!col_alive(col)) continue;
            A[size_t(w++)] = col;
            A[size_t(cstart[col] + clen[col]++)] = prow;
            const int64_t max_score = n_col - k - thick_or_parent[col];
            const int64_t cur = std::min(score_or_order[col] + pdeg - thick_or_parent[col], max_score);
            score_or_order[col] = cur;
            const int64_t nx = head[cur];
            next_link[col] = nx;
            prev_or_hash[col] = kEmpty;
            if (nx != kEmpty) prev_or_hash[nx] = col;
            head[cur] = col;
            min_score = std::min(min_score, cur);
        }
        A.resize(size_t(w));  // the compacted pivot row stays; the store grows from its end
        if (pdeg > 0) {
            rstart[prow] = prs;
            rlen[prow] = w - prs;
            degree[prow] = pdeg;
            mark[prow] = 0;
        }
    }
    // absorbed columns take the places right after their principal column
    for (int i = 0; i < n_col; i++) {
        if (cstart[i] == kDeadPrincipal || score_or_order[i] != kEmpty) continue;
        int64_t parent = i;
        do parent = thick_or_parent[parent]; while (cstart[parent] != kDeadPrincipal);
        int64_t c = i, ord = score_or_order[parent];
        do {
            score_or_order[c] = ord++;
            thick_or_parent[c] = parent;
            c = thick_or_parent[c];
        } while (score_or_order[c] == kEmpty);
        score_or_order[parent] = ord;
    }
    std::vector<int> p(static_cast<size_t>(n_col));
    for (int c = 0; c < n_col; c++) p[size_t(score_or_order[c])] = c;
    return p;
}

}  // namespace colamd
}  // namespace metriko

#endif
