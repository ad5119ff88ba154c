/* Test problem for the project build: the external potential of the
 * td-gauss-slda testsuite case, written against the reference's callback
 * interface (td-project-template/problem-definition.h).
 *
 * params (after process_params): 0 amplitude (energy), 1..3 exponent
 * coefficients -1/(2 s^2) along x, y, z (0: flat), 4..6 switch-on time,
 * switch-off time and switching time, 7 > 0.5: opposite sign for spin b.
 * V_ext = -A s(t) exp(w_x x^2 + w_y y^2 + w_z z^2), s(t) the smooth step
 * 1/2 (1 + tanh(tan(pi/2 (2 t / T - 1)))) up over [t_on, t_on + T] and down
 * over [t_off - T, t_off]. */

__device__ __host__ inline double ramp(double t, double T)
{
    return 0.5 * (1.0 + tanh(tan(0.5 * M_PI * (2.0 * t / T - 1.0))));
}

__device__ __host__ inline double window(double t, double on, double off, double T)
{
    if (t <= on || t >= off) return 0.0;
    if (t < on + T) return ramp(t - on, T);
    if (t > off - T) return 1.0 - ramp(t - off + T, T);
    return 1.0;
}

extern "C" void process_params(double *params, double *kF, double *mu, size_t extra_data_size, void *extra_data)
{
    const double eF = 0.5 * kF[0] * kF[0];
    params[0] *= eF;                         /* amplitude in eF            */
    for (int d = 1; d <= 3; d++) {           /* widths in 1/kF             */
        const double s = params[d] / kF[0];
        params[d] = s > 0.1 ? -0.5 / (s * s) : 0.0;
    }
    for (int k = 4; k <= 6; k++) params[k] /= eF;   /* times in 1/eF       */
}

__device__ double v_ext(int ix, int iy, int iz, int it, int spin, double *params, size_t extra_data_size,
                        void *extra_data)
{
    const double x = DX * (ix - NX / 2), y = DY * (iy - NY / 2), z = DZ * (iz - NZ / 2);
    const double t = dc_t0 + dc_dt * it;
    double v = -params[0] * window(t, params[4], params[5], params[6]) *
               exp(params[1] * x * x + params[2] * y * y + params[3] * z * z);
    if (spin == SPINB && params[7] > 0.5) v = -v;
    return v;
}

__device__ Complex delta_ext(int ix, int iy, int iz, int it, Complex delta, double *params, size_t extra_data_size,
                             void *extra_data)
{
    return Complex(0.0, 0.0);
}

__device__ double velocity_ext(int ix, int iy, int iz, int it, int spin, int coordinate, double *params,
                               size_t extra_data_size, void *extra_data)
{
    return 0.0;
}

extern "C" size_t get_extra_data_size(double *params) { return 0; }

extern "C" int load_extra_data(size_t extra_data_size, void *extra_data, double *params) { return 0; }

__device__ double scattering_length(int ix, int iy, int iz, int it, double *params, size_t extra_data_size,
                                    void *extra_data)
{
    return dc_sclgth;
}
