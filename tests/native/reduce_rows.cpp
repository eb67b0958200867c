// Host driver for the independent-row reduction of the drop-in headers
// (include/metriko/core/solver/matrix_ops.h, colamd.h): reads matrices in CSR form from stdin
// ("nr nc nnz", then the row pointers, column indices and values) and prints, per matrix, the
// COLAMD column order of C^T and the rows reduce_to_linearly_independent keeps, in order.
#include <cstdio>
#include <vector>
#include "metriko/core/solver/matrix_ops.h"

int main() {
    int nr, nc;
    long long nnz;
    while (std::scanf("%d %d %lld", &nr, &nc, &nnz) == 3) {
        metriko::SprsD C(nr, nc);
        C.ptr.resize(size_t(nr) + 1);
        C.col.resize(size_t(nnz));
        C.val.resize(size_t(nnz));
        for (int i = 0; i <= nr; i++) { long long v; if (std::scanf("%lld", &v) != 1) return 2; C.ptr[i] = v; }
        for (long long k = 0; k < nnz; k++) if (std::scanf("%d", &C.col[size_t(k)]) != 1) return 2;
        for (long long k = 0; k < nnz; k++) if (std::scanf("%lf", &C.val[size_t(k)]) != 1) return 2;
        const std::vector<int> p = metriko::colamd::order(nc, nr, C.ptr, C.col);
        std::printf("order");
        for (int x : p) std::printf(" %d", x);
        std::printf("\n");
        // the kept rows: tag each row by its index in a spare column-free copy of the values
        metriko::SprsD R = C;
        std::vector<int> before(size_t(nr));
        metriko::reduce_to_linearly_independent(R);
        // identify the kept rows by matching the row slices (the reduction keeps them verbatim)
        std::printf("kept");
        int from = 0;
        for (int i = 0; i < R.rows(); i++) {
            int found = -1;
            for (int r = 0; r < nr && found < 0; r++) {
                if (C.ptr[r + 1] - C.ptr[r] != R.ptr[i + 1] - R.ptr[i]) continue;
                bool same = true;
                for (int64_t k = 0; k < C.ptr[r + 1] - C.ptr[r] && same; k++)
                    same = C.col[C.ptr[r] + k] == R.col[R.ptr[i] + k] && C.val[C.ptr[r] + k] == R.val[R.ptr[i] + k];
                if (same) found = r;
            }
            std::printf(" %d", found);
            (void)from;
        }
        std::printf("\n");
        std::fflush(stdout);
    }
    return 0;
}
