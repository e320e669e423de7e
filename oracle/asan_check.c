/* Host sanitizer check of the CPU oracle (test infrastructure): built with
 * -fsanitize=address,undefined, it runs every oracle entry point on small lattices --
 * Langevin steps (plain and even-odd force, both step-size modes), the coloured
 * filter on odd FFT lengths, CG in both operator orders, the measurement correlator,
 * the CNN forward pass and loss gradient -- and exits non-zero on the first sanitizer
 * report (halt_on_error) or a non-converged solve. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ycn_oracle.h"

static orc_params params(int Nt, int Nx) {
    orc_params p;
    memset(&p, 0, sizeof p);
    p.Nt = Nt; p.Nx = Nx; p.m2 = -0.4; p.lambda = 1.0; p.g = 0.5; p.M = 0.05; p.r = 1.0;
    p.cutFraction = 0.5; p.epsBar = 0.005; p.seed = 1234; p.cgTol = 1e-10; p.cgTolMode = 1;
    p.cgMaxIter = 20000; p.nNoise = 1;
    return p;
}

int main(void) {
    int bad = 0;
    const int shapes[][2] = {{8, 8}, {6, 10}, {2, 12}, {16, 16}};
    for (size_t k = 0; k < sizeof shapes / sizeof shapes[0]; ++k) {
        const int Nt = shapes[k][0], Nx = shapes[k][1], V = Nt * Nx;
        double *phi = malloc(sizeof(double) * V), *xi = malloc(sizeof(double) * V);
        double *out = malloc(sizeof(double) * V), *corr = malloc(sizeof(double) * Nt);
        double *b = malloc(sizeof(double) * 8 * V), *x = malloc(sizeof(double) * 8 * V);
        orc_cg_stats st;
        for (int eo = 0; eo < 2; ++eo)
            for (int adaptive = 0; adaptive < 2; ++adaptive) {
                orc_params p = params(Nt, Nx);
                p.evenOdd = eo;
                p.adaptive = adaptive;
                orc_hot_start(7, V, phi);
                for (int step = 0; step < 3; ++step) {
                    double eps = 0;
                    orc_langevin_step(&p, phi, (uint64_t)step, &eps, &st);
                    bad |= !st.converged;
                }
            }
        orc_params p = params(Nt, Nx);
        orc_normals(3, 0, 0, V, xi);
        orc_coloured_filter(Nt, Nx, 0.25, xi, out);
        orc_fermion_noise(p.seed, 0, 0, V, b);
        for (int mtype = 0; mtype < 2; ++mtype) {
            orc_cg(&p, phi, b, x, mtype, 0, &st);
            bad |= !st.converged;
        }
        orc_correlator(&p, phi, corr, &st);
        if (Nt % 2 == 0 && Nx % 2 == 0) {
            orc_eo_cg(&p, phi, b, x, 0, &st);
            bad |= !st.converged;
        }
        float *W = calloc(ORC_CNN_NW, sizeof(float));
        double *grad = malloc(sizeof(double) * ORC_CNN_NW);
        for (int i = 0; i < ORC_CNN_NW; ++i) W[i] = 0.01f * (float)((i * 37) % 19 - 9);
        orc_cnn_forward(Nt, Nx, W, 1, 0.2, p.g, phi, b, x);
        orc_normal_op(&p, phi, b, x, 1);
        (void)orc_cnn_loss_grad(&p, phi, W, 0.2, x, b, grad);
        free(W); free(grad);
        free(phi); free(xi); free(out); free(corr); free(b); free(x);
    }
    printf("oracle sanitizer check: %s\n", bad ? "a solve did not converge" : "clean");
    return bad;
}
