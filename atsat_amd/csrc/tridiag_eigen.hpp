// Eigen-decomposition of a symmetric tridiagonal matrix by the implicit QL method with Wilkinson
// shifts (Golub & Van Loan, "Matrix Computations", section 8.3; Wilkinson & Reinsch, "Handbook
// for Automatic Computation II", contribution II/3). Host code, used once per context to set up
// the separable pressure solver. Header-only so that it can also be compiled on its own.
#pragma once

#include <cmath>
#include <vector>

namespace atsat {

// d[n]: the diagonal (overwritten by the eigenvalues); e[n]: e[i] = T(i, i+1) for i < n-1
// (destroyed); z[n*n]: on return column c (z[r*n + c]) is the unit eigenvector of d[c].
// Returns false if an eigenvalue did not converge in 60 iterations.
inline bool tridiag_eigen(int n, std::vector<double>& d, std::vector<double> e, std::vector<double>& z){
    z.assign((size_t)n * n, 0.0);
    for(int i = 0; i < n; i++) z[(size_t)i * n + i] = 1.0;
    if(n == 0) return true;
    e.resize(n);
    e[n - 1] = 0.0;
    const double eps = 1e-16;
    for(int l = 0; l < n; l++){
        int iter = 0;
        while(true){
            int m = l;
            for(; m < n - 1; m++){
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if(std::fabs(e[m]) <= eps * dd) break;
            }
            if(m == l) break;
            if(++iter > 60) return false;
            // shift from the trailing 2x2 block of the unreduced part starting at l
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? r : -r));
            double s = 1.0, c = 1.0, p = 0.0;
            bool deflated = false;
            for(int i = m - 1; i >= l; i--){
                const double f = s * e[i], b = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if(r == 0.0){                         // underflow: split the problem
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    deflated = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                for(int k = 0; k < n; k++){           // accumulate the rotation
                    double* zk = &z[(size_t)k * n];
                    const double t = zk[i + 1];
                    zk[i + 1] = s * zk[i] + c * t;
                    zk[i] = c * zk[i] - s * t;
                }
            }
            if(deflated) continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    return true;
}

}  // namespace atsat
