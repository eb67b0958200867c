// Host driver for the FullPivLU kernel of the drop-in headers
// (include/metriko/core/solver/matrix_ops.h, fullpiv_lu_kernel): reads dense matrices from
// stdin ("rows cols", then the entries row by row) and prints the rank and the kernel basis
// (cols x max(1, dimker), row by row, 17 significant digits).
#include <cstdio>
#include "metriko/core/solver/matrix_ops.h"

int main() {
    int r, c;
    while (std::scanf("%d %d", &r, &c) == 2) {
        metriko::MatXd A(r, c);
        for (int i = 0; i < r; i++)
            for (int j = 0; j < c; j++)
                if (std::scanf("%lf", &A(i, j)) != 1) return 2;
        int rank = 0;
        const metriko::MatXd K = metriko::fullpiv_lu_kernel(A, &rank);
        std::printf("rank %d %d %d\n", rank, K.rows(), K.cols());
        for (int i = 0; i < K.rows(); i++) {
            for (int j = 0; j < K.cols(); j++) std::printf("%.17g ", K(i, j));
            std::printf("\n");
        }
        std::fflush(stdout);
    }
    return 0;
}
