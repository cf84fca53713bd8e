/* A C99 caller of libcutcell_hip that uses nothing but include/cutcell_hip.h -- what a Julia
 * HIPBackend binding does through ccall.  Built and run by tests/test_abi_c.py.
 *
 *   abi_client version            prints cc_version() (needs no GPU)
 *   abi_client run OUT.bin        on the first GPU: allocate_cache / update_cache! / readers /
 *                                 error paths for a 3D marching-cubes cache, then writes the kind,
 *                                 volume_fraction, face_fraction and phi columns to OUT.bin
 *
 * Every status is checked; the first failure prints cc_last_error() and exits non-zero. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cutcell_hip.h"

#define CHECK(call)                                                                            \
    do {                                                                                       \
        int rc_ = (call);                                                                      \
        if (rc_ != CC_OK) {                                                                    \
            fprintf(stderr, "%s:%d: %s -> %d: %s\n", __FILE__, __LINE__, #call, rc_, cc_last_error()); \
            exit(1);                                                                           \
        }                                                                                      \
    } while (0)

static void fail(const char *what) {
    fprintf(stderr, "abi_client: %s\n", what);
    exit(1);
}

static void *read_column(cc_cache *c, int field, int64_t *nbytes) {
    void *dev = NULL;
    CHECK(cc_column(c, field, &dev, nbytes));
    if (dev == NULL || *nbytes <= 0) fail("a column without storage");
    void *host = malloc((size_t)*nbytes);
    if (!host) fail("out of host memory");
    CHECK(cc_copy_to_host(c, field, 0, *nbytes, host));
    return host;
}

int main(int argc, char **argv) {
    if (argc >= 2 && strcmp(argv[1], "version") == 0) {
        printf("%s\n", cc_version());
        return 0;
    }
    if (argc < 3 || strcmp(argv[1], "run") != 0) fail("usage: abi_client version | run OUT.bin");

    /* the grid and body of the smoke check: 24^3 cells of 0.125 from -1.5, a unit sphere */
    cc_grid g;
    memset(&g, 0, sizeof g);
    g.dim = 3;
    for (int a = 0; a < 3; ++a) {
        g.n[a] = 24;
        g.x0[a] = -1.5;
        g.d[a] = 0.125;
    }
    cc_geometry sphere;
    memset(&sphere, 0, sizeof sphere);
    sphere.kind = CC_GEO_SPHERE;
    sphere.p[0] = 0.01;
    sphere.p[1] = 0.02;
    sphere.p[2] = -0.03;
    sphere.p[3] = 1.0;

    cc_cache *c = NULL;
    CHECK(cc_cache_create(CC_MARCHING_CUBES, CC_FLOAT64, &g, &c));
    CHECK(cc_update(c, &sphere, &g));

    const int64_t ncell = g.n[0] * g.n[1] * g.n[2];
    int64_t nb_kind, nb_vf, nb_ff, nb_phi, ncut = -1;
    int8_t *kind = (int8_t *)read_column(c, CC_FIELD_KIND, &nb_kind);
    double *vf = (double *)read_column(c, CC_FIELD_VOLUME_FRACTION, &nb_vf);
    double *ff = (double *)read_column(c, CC_FIELD_FACE_FRACTION, &nb_ff);
    double *phi = (double *)read_column(c, CC_FIELD_PHI, &nb_phi);
    if (nb_kind != ncell || nb_vf != 8 * ncell || nb_ff != 6 * 8 * ncell || nb_phi != 8 * 25 * 25 * 25)
        fail("column sizes differ from the layout the header states");

    /* the readers' invariants: kinds against fractions, the cut-cell count */
    int64_t counted = 0;
    for (int64_t i = 0; i < ncell; ++i) {
        if (kind[i] == CC_CELL_CUT) ++counted;
        else if (kind[i] == CC_CELL_INSIDE && vf[i] != 0.0) fail("an inside cell with fluid");
        else if (kind[i] == CC_CELL_OUTSIDE && vf[i] != 1.0) fail("an outside cell without fluid");
        if (!(vf[i] >= 0.0 && vf[i] <= 1.0)) fail("a volume fraction outside [0, 1]");
    }
    CHECK(cc_cut_count(c, &ncut));
    if (ncut != counted || counted < 1000) fail("cc_cut_count disagrees with the kind column");

    /* update_cache! with a grid of another size: DimensionMismatch, with a message */
    cc_grid wrong = g;
    wrong.n[0] = 10;
    if (cc_update(c, &sphere, &wrong) != CC_ERR_DIMENSION_MISMATCH) fail("a wrong n was accepted");
    if (strlen(cc_last_error()) == 0) fail("no message for the failed call");
    /* an anisotropic grid for PLIC: ArgumentError */
    cc_grid aniso = g;
    aniso.d[1] = 0.25;
    cc_cache *p = NULL;
    if (cc_cache_create(CC_PLIC, CC_FLOAT64, &aniso, &p) != CC_ERR_ARGUMENT || p != NULL)
        fail("an anisotropic PLIC grid was accepted");

    /* the cache still answers after the refused calls */
    CHECK(cc_update(c, &sphere, &g));

    FILE *f = fopen(argv[2], "wb");
    if (!f) fail("cannot open the output file");
    if (fwrite(kind, 1, (size_t)nb_kind, f) != (size_t)nb_kind || fwrite(vf, 1, (size_t)nb_vf, f) != (size_t)nb_vf ||
        fwrite(ff, 1, (size_t)nb_ff, f) != (size_t)nb_ff || fwrite(phi, 1, (size_t)nb_phi, f) != (size_t)nb_phi)
        fail("short write");
    fclose(f);
    free(kind);
    free(vf);
    free(ff);
    free(phi);
    CHECK(cc_cache_destroy(c));
    printf("abi_client: %lld cells, %lld cut\n", (long long)ncell, (long long)ncut);
    return 0;
}
