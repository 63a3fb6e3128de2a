// Host-side grid metrics of lib/grid/grid.cc (createUniformGrid 370-422, createTanHypGrid 442-519):
// node ("staggered") and face ("collocated") coordinates of one direction with the analytic
// metric terms xi_x, xi_xx and xi_x^2 that every stencil of the solver reads.
#include <cmath>

#include "../../include/saras_mg.h"

extern "C" int sr_grid_axis(int n, double len, int grid_type, double beta, int faces, double *x, double *xi_x,
                            double *xixx, double *xix2) {
    if (n < 3 || len <= 0.0 || !x || !xi_x || !xixx || !xix2 || (grid_type != 1 && grid_type != 2)) return SR_ERR_ARG;
    if (grid_type == 2 && beta <= 0.0) return SR_ERR_ARG;
    const double dxi = 1.0 / (double)(n - 1);
    // uniform values everywhere, pads included (grid.cc:373-420)
    const int last = faces ? n - 1 : n;
    for (int i = -1; i <= last; i++) {
        double a = (double)i * dxi;
        x[i + 1] = faces ? len * (a + (double)(i + 1) * dxi) / 2.0 : len * a;
        xi_x[i + 1] = 1.0 / len;
        xixx[i + 1] = 0.0;
        xix2[i + 1] = std::pow(1.0 / len, 2.0);
    }
    if (grid_type == 2) {
        // tanh stretching overwrites the core indices only (grid.cc:461-477)
        const double tb = std::tanh(beta);
        const int core = faces ? n - 1 : n;
        for (int i = 0; i < core; i++) {
            double arg = faces ? 1.0 - ((double)i * dxi + (double)(i + 1) * dxi) : 1.0 - 2.0 * (double)i * dxi;
            double xs = len * (1.0 - std::tanh(beta * arg) / tb) / 2.0;
            double s = 1.0 - 2.0 * xs / len;
            x[i + 1] = xs;
            xi_x[i + 1] = tb / (beta * len * (1.0 - std::pow(s * tb, 2)));
            xixx[i + 1] = -4.0 * std::pow(tb, 3) * s / (beta * len * len * std::pow(1.0 - std::pow(tb * s, 2), 2));
            xix2[i + 1] = std::pow(xi_x[i + 1], 2.0);
        }
    }
    return SR_OK;
}
