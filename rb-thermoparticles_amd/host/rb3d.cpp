// rb3d — host driver with the reference's time loop (main.cu:4-271) on top of the rbtp C-ABI.
//
//   rb3d [parameters.h] [--precision fp64|fp32|fp16s|bf16s] [--temperature linear:SEED:AMP]
//        [--nend N]
//
// The case comes from a parameters.h file (same macro names as the reference; defaults are the
// reference's own values).  Output cadence, file names and formats follow main.cu:98-114:
// statistics.dat every NDIAG steps, %09d.dat every NFLOWOUT steps and at the end,
// pos/pos_%09d.txt every NOUT_P steps, Nu_%.2e.txt and Profile_%.2e.txt every NNUOUT steps,
// fg%09d.dat at the end, all in the directory Ra%.1ePr%.2f.  CONTINUOUS_STEPS > 0 restarts from
// that directory (initial.cu:194-236).  Steps between outputs are enqueued back to back; the GPU
// time reported is the device time of those steps (hipEvents on the solver's stream).
#include <sys/stat.h>
#include <sys/time.h>
#include <sys/types.h>

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "rbtp.h"

namespace {

void die(const char *what) {
    fprintf(stderr, "%s: %s\n", what, rbtp_last_error());
    exit(EXIT_FAILURE);
}

#define CHECK(call)              \
    do {                         \
        if ((call) != RBTP_OK) die(#call); \
    } while (0)

void ensure_dir(const std::string &d) {
    struct stat st;
    if (stat(d.c_str(), &st) == -1 && mkdir(d.c_str(), 0700) != 0 && errno != EEXIST) {
        fprintf(stderr, "cannot create %s\n", d.c_str());
        exit(EXIT_FAILURE);
    }
}

struct Driver {
    rbtp_case c;
    rbtp_derived d;
    rbtp_sim *sim = nullptr;
    std::string dir;
    bool particles = false;

    // diag_flow (output_statis.cu:64-109)
    void diag_flow(int istep) {
        ensure_dir(dir);
        double m[4];
        CHECK(rbtp_means(sim, m));
        printf("Step %-7d umean = %14.6e, vmean = %14.6e, wmean = %14.6e, tmean = %14.6e\n", istep,
               m[0], m[1], m[2], m[3]);
        fflush(stdout);
        FILE *f = fopen((dir + "/statistics.dat").c_str(), "a");
        if (!f) die("statistics.dat");
        fprintf(f, "%-7d %14.6e %14.6e %14.6e %14.6e\n", istep, m[0], m[1], m[2], m[3]);
        fclose(f);
    }
    // output_flow (output_flow.cu:64-110)
    void output_flow(int istep) {
        ensure_dir(dir);
        char name[512];
        snprintf(name, sizeof(name), "%s/%09d.dat", dir.c_str(), istep);
        CHECK(rbtp_write_flow(sim, name));
    }
    // output_particles (output_part.cu:4-119)
    void output_particles(int istep) {
        if (!particles) return;
        ensure_dir(dir);
        ensure_dir(dir + "/pos");
        char name[1024];
        snprintf(name, sizeof(name), "%s/pos/pos_%09d.txt", dir.c_str(), istep);
        CHECK(rbtp_write_particles(sim, name));
    }
    // output_nu (output_statis.cu:111-223)
    void output_nu(int istep) {
        ensure_dir(dir);
        double nu[7];
        CHECK(rbtp_nu(sim, nu));
        char name[512];
        snprintf(name, sizeof(name), "%s/Nu_%.2e.txt", dir.c_str(), c.rayleigh);
        FILE *f = fopen(name, "a");
        if (!f) die("Nu file");
        fprintf(f, "%15.6E%15.6E%15.6E%15.6E%15.6E%15.6E%15.6E\n", (double)istep, nu[1], nu[2],
                nu[3], nu[4], nu[5], nu[6]);
        fclose(f);
    }
    // output_profile (output_statis.cu:225-285)
    void output_profile(int istep) {
        ensure_dir(dir);
        std::vector<double> prof(4 * (size_t)c.lx);
        CHECK(rbtp_profile(sim, prof.data()));
        char name[512];
        snprintf(name, sizeof(name), "%s/Profile_%.2e.txt", dir.c_str(), c.rayleigh);
        FILE *f = fopen(name, "a");
        if (!f) die("Profile file");
        fprintf(f, "%d\n", istep);
        for (int ix = 0; ix < c.lx; ix++)
            fprintf(f, "%14.6e %14.6e %14.6e %14.6e\n", prof[4 * ix], prof[4 * ix + 1],
                    prof[4 * ix + 2], prof[4 * ix + 3]);
        fclose(f);
    }
    // output_fg (output_flow.cu:112-127)
    void output_fg(int istep) {
        ensure_dir(dir);
        char name[512];
        snprintf(name, sizeof(name), "%s/fg%09d.dat", dir.c_str(), istep);
        CHECK(rbtp_write_fg(sim, name));
    }
    bool is_output_step(int istep, int final_step) const {
        return istep % c.ndiag == 0 || istep % c.nflowout == 0 || istep == final_step ||
               istep % c.nout_p == 0 || istep % c.nnuout == 0;
    }
};

int parse_precision(const char *s) {
    if (!strcmp(s, "fp64")) return RBTP_PREC_FP64;
    if (!strcmp(s, "fp32")) return RBTP_PREC_FP32;
    if (!strcmp(s, "fp16s")) return RBTP_PREC_FP16S;
    if (!strcmp(s, "bf16s")) return RBTP_PREC_BF16S;
    fprintf(stderr, "unknown precision %s\n", s);
    exit(EXIT_FAILURE);
}

}  // namespace

int main(int argc, char **argv) {
    struct timeval start_time, end_time;
    gettimeofday(&start_time, nullptr);
    Driver D;
    CHECK(rbtp_case_default(&D.c));
    const char *temp_spec = nullptr;
    int precision = -1, nend = -1;
    for (int a = 1; a < argc; a++) {
        if (!strcmp(argv[a], "--precision") && a + 1 < argc) precision = parse_precision(argv[++a]);
        else if (!strcmp(argv[a], "--temperature") && a + 1 < argc) temp_spec = argv[++a];
        else if (!strcmp(argv[a], "--nend") && a + 1 < argc) nend = atoi(argv[++a]);
        else if (argv[a][0] != '-') CHECK(rbtp_case_from_parameters_h(argv[a], &D.c));
        else {
            fprintf(stderr, "usage: %s [parameters.h] [--precision P] [--temperature linear:SEED:AMP] [--nend N]\n", argv[0]);
            return EXIT_FAILURE;
        }
    }
    if (precision >= 0) D.c.precision = precision;
    if (nend >= 0) D.c.nend = nend;
    CHECK(rbtp_case_derive(&D.c, &D.d));
    D.particles = D.c.activate_particles && D.d.npart > 0;
    char dirname[256];
    snprintf(dirname, sizeof(dirname), "Ra%.1ePr%.2f", D.c.rayleigh, D.c.prandtl);
    D.dir = dirname;
    CHECK(rbtp_create(&D.c, &D.sim));
    const int start_step = D.c.continuous_steps > 0 ? D.c.continuous_steps : 0;
    if (start_step > 0) {
        CHECK(rbtp_restart_load(D.sim, D.dir.c_str(), start_step));
    } else {
        std::vector<double> T;
        if (temp_spec) {
            unsigned long long seed = 0;
            double amp = 0.0;
            if (sscanf(temp_spec, "linear:%llu:%lf", &seed, &amp) != 2) {
                fprintf(stderr, "--temperature expects linear:SEED:AMP\n");
                return EXIT_FAILURE;
            }
            T.resize((size_t)D.c.lx * D.c.ly * D.c.lz);
            CHECK(rbtp_linear_temperature(&D.c, seed, amp, T.data()));
        }
        CHECK(rbtp_init(D.sim, T.empty() ? nullptr : T.data(), nullptr, nullptr));
    }
    const int final_step = D.c.nend;
    if (D.particles && start_step == 0) D.output_particles(0);

    double gpu_time_used = 0.0, cpu_time_used = 0.0;
    int istep = start_step;
    while (istep < final_step) {
        // run to the next step that writes something (main.cu:98-114)
        int next = istep + 1;
        while (next < final_step && !D.is_output_step(next, final_step)) next++;
        CHECK(rbtp_step(D.sim, next - istep, RBTP_STEP_MACRO_LAST));
        double ms = 0.0;
        CHECK(rbtp_last_step_ms(D.sim, &ms));
        gpu_time_used += ms / 1000.0;
        istep = next;
        const clock_t cpu_start = clock();
        if (istep % D.c.ndiag == 0) D.diag_flow(istep);
        if (istep % D.c.nflowout == 0 || istep == final_step) D.output_flow(istep);
        if (istep % D.c.nout_p == 0) D.output_particles(istep);
        if (istep % D.c.nnuout == 0) {
            D.output_nu(istep);
            D.output_profile(istep);
        }
        if (istep == final_step) {
            D.output_fg(istep);
            D.output_particles(istep);
        }
        cpu_time_used += (double)(clock() - cpu_start) / CLOCKS_PER_SEC;
    }
    gettimeofday(&end_time, nullptr);
    const double total_time = (end_time.tv_sec - start_time.tv_sec) +
                              (end_time.tv_usec - start_time.tv_usec) / 1000000.0;
    printf("Simulation completed.\n");
    printf("Rayleigh number: %.2e\n", D.c.rayleigh);
    printf("Prandtl number: %.2f\n", D.c.prandtl);
    printf("Total iterations: %d\n", D.c.nend);
    if (start_step > 0) printf("Restarted from step: %d, final step: %d\n", start_step, final_step);
    else printf("Final step: %d\n", final_step);
    printf("CPU computation time: %.2f seconds\n", cpu_time_used);
    printf("GPU computation time: %.2f seconds\n", gpu_time_used);
    printf("Total computation time: %.2f seconds\n", total_time);
    const double nodes = (double)D.c.lx * D.c.ly * D.c.lz;
    if (gpu_time_used > 0)
        printf("GPU throughput: %.1f MLUPS\n", nodes * (final_step - start_step) / gpu_time_used / 1e6);
    rbtp_destroy(D.sim);
    return 0;
}
