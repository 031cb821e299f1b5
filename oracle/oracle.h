/* C declarations of the CPU oracle (test infrastructure; see alsvinn_oracle.cpp). */
#ifndef ALSVINN_ORACLE_H
#define ALSVINN_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_setup {
    int equation, flux, recon, integrator, boundary;
    int nx, ny, nz;
    double origin[3], top[3];
    double gamma, cfl, end_time;
    int num_adjusters;
    double adj_interval[2];
    long adj_saved0[2];
    long max_steps;
} oracle_setup;

int oracle_ghost_cells(const oracle_setup* s);
int oracle_run(const oracle_setup* setup, const double* u0, double* uout, double* t_out,
               long* steps_out, double* dt_hist, long dt_cap);
int oracle_system_rhs(const oracle_setup* setup, const double* u, double* out, double* ws3);
int oracle_compute_flux(const oracle_setup* setup, const double* u, const int* start,
                        const int* end, double* out, double* ws3);
int oracle_reconstruct(const oracle_setup* setup, const double* u, int direction, const int* start,
                       const int* end, double* left, double* right);
int oracle_apply_boundary(const oracle_setup* setup, double* u);
void oracle_meanvar(long nsamples, long len, const double* samples, double* mean, double* var);
void oracle_onepoint_moment(long nsamples, long len, const double* samples, double p,
                            double* out);
double oracle_total_variation(int tx, int ty, int tz, int gx, int gy, int gz, const double* u,
                              double p);
void oracle_structure_basic(int nvar, int nx, int ny, int nz, int ng, const double* u, int H,
                            double p, int direction, int periodic, double* out);
void oracle_stl_mersenne_uniform(long sample_start, long nsamples, long dim, double lower,
                                 double upper, double* out);

#ifdef __cplusplus
}
#endif
#endif
