/*
 * ref_runner.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Drives the reference HDSDP solver (compiled from /root/reference by
 * oracle/Makefile.ref into oracle/_ref/) on one instance file and prints its
 * iteration log followed by one machine-readable RESULT line.  Used to write
 * golden fixtures and as the CPU baseline; never linked into the product.
 *
 * Instance file layout: see hdsdp_amd/instances.py (write_instance).
 *
 * usage: hdsdp_ref_run INSTANCE YOUT [key=value ...]
 *   psdp=0|1 relopt=X absopt=X maxiter=K timelimit=T threads=K corra=K corrb=K
 *
 * The public setters reject the PSDP switch and the thread count
 * (interface/hdsdp.c:596-604: indices 3 and 4 are dropped), so both are
 * written through the solver struct (interface/def_hdsdp.h:131-134) after
 * HDSDPInit, which the solver never resets (interface/hdsdp.c:130-376).  The
 * pre-solve then caps the thread count at what MKL reports
 * (interface/hdsdp.c:306-315) and prints the count it uses ("Optimizing over N
 * thread(s)", interface/hdsdp.c:669).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#include <sys/time.h>

#include "def_hdsdp.h"
#include "hdsdp.h"
#include "hdsdp_user_data.h"

static double now_s(void) {
    struct timeval t;
    gettimeofday(&t, NULL);
    return t.tv_sec + 1e-6 * t.tv_usec;
}

static void *xread(FILE *f, size_t bytes) {
    void *p = malloc(bytes ? bytes : 1);
    if (!p || fread(p, 1, bytes, f) != bytes) {
        fprintf(stderr, "ref_runner: short read\n");
        exit(3);
    }
    return p;
}

int main(int argc, char **argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s INSTANCE YOUT [key=value ...]\n", argv[0]);
        return 2;
    }
    int psdp = 1, maxiter = -1, corra = -1, corrb = -1, threads = -1;
    double relopt = -1.0, absopt = -1.0, timelimit = -1.0;
    for (int a = 3; a < argc; ++a) {
        if (!strncmp(argv[a], "psdp=", 5)) psdp = atoi(argv[a] + 5);
        else if (!strncmp(argv[a], "relopt=", 7)) relopt = atof(argv[a] + 7);
        else if (!strncmp(argv[a], "absopt=", 7)) absopt = atof(argv[a] + 7);
        else if (!strncmp(argv[a], "maxiter=", 8)) maxiter = atoi(argv[a] + 8);
        else if (!strncmp(argv[a], "timelimit=", 10)) timelimit = atof(argv[a] + 10);
        else if (!strncmp(argv[a], "corra=", 6)) corra = atoi(argv[a] + 6);
        else if (!strncmp(argv[a], "corrb=", 6)) corrb = atoi(argv[a] + 6);
        else if (!strncmp(argv[a], "threads=", 8)) threads = atoi(argv[a] + 8);
        else { fprintf(stderr, "unknown option %s\n", argv[a]); return 2; }
    }

    FILE *f = fopen(argv[1], "rb");
    if (!f) { perror("instance"); return 3; }
    char magic[8];
    if (fread(magic, 1, 8, f) != 8 || memcmp(magic, "HDSDPIN1", 8)) { fprintf(stderr, "bad magic\n"); return 3; }
    int32_t hdr[2]; int64_t nnz;
    if (fread(hdr, 4, 2, f) != 2 || fread(&nnz, 8, 1, f) != 1) return 3;
    int n = hdr[0], m = hdr[1];
    if (nnz > INT32_MAX) { fprintf(stderr, "instance exceeds the int32 CSC interface\n"); return 4; }
    int64_t *beg64 = xread(f, sizeof(int64_t) * (m + 2));
    int *idx = xread(f, sizeof(int) * nnz);
    double *elem = xread(f, sizeof(double) * nnz);
    double *b = xread(f, sizeof(double) * m);
    fclose(f);
    int *beg = malloc(sizeof(int) * (m + 2));
    for (int k = 0; k < m + 2; ++k) beg[k] = (int) beg64[k];

    double t0 = now_s();
    hdsdp *solver = NULL;
    user_data *ud = NULL;
    int rc = 0;
    if ((rc = HUserDataCreate(&ud))) goto done;
    HUserDataSetConeData(ud, HDSDP_CONETYPE_DENSE_SDP, m, n, beg, idx, elem);
    if ((rc = HDSDPCreate(&solver))) goto done;
    if ((rc = HDSDPInit(solver, m, 1))) goto done;
    if ((rc = HDSDPSetCone(solver, 0, ud))) goto done;
    HDSDPSetDualObjective(solver, b);
    solver->HIntParams[INT_PARAM_PSDP] = psdp;
    if (relopt > 0) HDSDPSetDblParam(solver, DBL_PARAM_RELOPTTOL, relopt);
    if (absopt > 0) HDSDPSetDblParam(solver, DBL_PARAM_ABSOPTTOL, absopt);
    if (maxiter > 0) HDSDPSetIntParam(solver, INT_PARAM_MAXITER, maxiter);
    if (timelimit > 0) HDSDPSetDblParam(solver, DBL_PARAM_TIMELIMIT, timelimit);
    if (threads > 0) solver->HIntParams[INT_PARAM_THREADS] = threads;
    (void) corra; (void) corrb;
    rc = HDSDPOptimize(solver, 1);
    double t1 = now_s();
    {
        double pobj = 0, dobj = 0;
        double *y = calloc(m, sizeof(double));
        HDSDPGetRowDual(solver, &pobj, &dobj, y);
        FILE *fy = fopen(argv[2], "wb");
        if (fy) { fwrite(y, sizeof(double), m, fy); fclose(fy); }
        fflush(stdout);
        printf("\nRESULT {\"retcode\": %d, \"status\": %d, \"iters\": %d, \"pobj\": %.17g, \"dobj\": %.17g, "
               "\"mu\": %.17g, \"time\": %.6f, \"threads\": %d, \"dimacs\": [%.17g, %.17g, %.17g, %.17g, %.17g, %.17g]}\n",
               rc, (int) solver->HStatus, solver->nIterCount, pobj, dobj, solver->dBarrierMu, t1 - t0,
               solver->HIntParams[INT_PARAM_THREADS], solver->dErrs[0], solver->dErrs[1], solver->dErrs[2], solver->dErrs[3], solver->dErrs[4],
               solver->dErrs[5]);
        fflush(stdout);
        free(y);
    }
done:
    if (rc) printf("\nRESULT {\"retcode\": %d, \"status\": -1}\n", rc);
    HDSDPDestroy(&solver);
    HUserDataDestroy(&ud);
    free(beg); free(beg64); free(idx); free(elem); free(b);
    return 0;
}
