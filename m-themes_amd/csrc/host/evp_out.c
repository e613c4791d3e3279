/* evp_out.c -- the reference's on-disk field and texture dumps.
 *
 * .iout field files (WriteSigMPI/WriteEpsMPI/WriteEdotMPI/WriteElsMPI/
 * WriteDisgradMPI, src/io.c:283-476): for one field, six files
 * "<prefix>_<c>_S<step, 6 digits>.iout", c = 1..6 for the components
 * 11 22 33 23 13 12, each holding the N voxel values as native doubles in
 * voxel order (x-major, z fastest; MPI_File_write of one slab per rank,
 * offset rank * lsize, which on one rank is the whole grid).
 *
 * Texture record (WriteTextureMPI, src/io.c:478-541, read back by
 * src/tex2out.c:18-70): one file "<prefix>_S<step, 6 digits>.iout" of N
 * 48-byte records {double angle[3]; int coord[3]; int grain; int phase;
 * int buffer;} with the Bunge angles in degrees of TranMat transposed
 * (EulerToTransMatrix option 1, src/kinematics.c:187-217), the 1-based
 * voxel coordinates, the grain and phase IDs and a zero buffer flag.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "evp_host.h"

#define EVP_PI_REF 3.14159265359 /* src/evp.h:77 (Q8) */

typedef struct {
    double angle[3];
    int32_t coord[3];
    int32_t grain, phase, buffer;
} tex_record;

_Static_assert(sizeof(tex_record) == 48, "texture record layout of src/io.c:480-487");

void evp_mat_to_euler(const double a9[9], double ang[3])
{
    double a[3][3];
    for (int i = 0; i < 9; i++) a[i / 3][i % 3] = a9[i];
    double t1, ph = acos(a[2][2]), t2;
    if (fabs(a[2][2]) > 0.9999) {
        t2 = 0.;
        t1 = atan2(a[0][1], a[0][0]);
    } else {
        double sph = sin(ph);
        t1 = atan2(a[2][0] / sph, -1.0 * a[2][1] / sph);
        t2 = atan2(a[0][2] / sph, a[1][2] / sph);
    }
    ang[0] = t1 * 180. / EVP_PI_REF;
    ang[1] = ph * 180. / EVP_PI_REF;
    ang[2] = t2 * 180. / EVP_PI_REF;
}

int evp_write_iout6(const char *prefix, int step, long long n, const double *f6)
{
    if (!prefix || !f6 || n < 1) return EVPGPU_ERR_ARG;
    char name[1024];
    for (int c = 0; c < 6; c++) {
        snprintf(name, sizeof(name), "%s_%01d_S%06d.iout", prefix, c + 1, step);
        FILE *fp = fopen(name, "wb");
        if (!fp) return EVP_ERR_FILE;
        size_t w = fwrite(f6 + (size_t)c * n, sizeof(double), (size_t)n, fp);
        if (fclose(fp) != 0 || w != (size_t)n) return EVP_ERR_FILE;
    }
    return 0;
}

int evp_write_texture(const char *prefix, int step, const int dims[3], const double *tran9,
                      const int *grain, const int *phase)
{
    if (!prefix || !dims || !tran9 || !grain || !phase) return EVPGPU_ERR_ARG;
    const long long n = (long long)dims[0] * dims[1] * dims[2];
    char name[1024];
    snprintf(name, sizeof(name), "%s_S%06d.iout", prefix, step);
    FILE *fp = fopen(name, "wb");
    if (!fp) return EVP_ERR_FILE;
    tex_record r;
    memset(&r, 0, sizeof(r));
    for (long long v = 0; v < n; v++) {
        double sa2xt[9];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) sa2xt[i * 3 + j] = tran9[(j * 3 + i) * n + v];
        evp_mat_to_euler(sa2xt, r.angle);
        r.coord[0] = (int32_t)(v / ((long long)dims[1] * dims[2])) + 1;
        r.coord[1] = (int32_t)((v / dims[2]) % dims[1]) + 1;
        r.coord[2] = (int32_t)(v % dims[2]) + 1;
        r.grain = grain[v];
        r.phase = phase[v];
        r.buffer = 0;
        if (fwrite(&r, sizeof(r), 1, fp) != 1) {
            fclose(fp);
            return EVP_ERR_FILE;
        }
    }
    return fclose(fp) == 0 ? 0 : EVP_ERR_FILE;
}

/* text listing of a texture record file, as src/tex2out.c:56-64 prints it */
int evp_texture_to_text(const char *in, const char *out)
{
    FILE *fi = fopen(in, "rb");
    if (!fi) return EVP_ERR_FILE;
    FILE *fo = fopen(out, "w");
    if (!fo) {
        fclose(fi);
        return EVP_ERR_FILE;
    }
    tex_record r;
    while (fread(&r, sizeof(r), 1, fi) == 1)
        fprintf(fo, "%lf\t%lf\t%lf\t%d\t%d\t%d\t%d\t%d\n", r.angle[0], r.angle[1], r.angle[2], r.coord[0],
                r.coord[1], r.coord[2], r.grain, r.phase);
    fclose(fi);
    return fclose(fo) == 0 ? 0 : EVP_ERR_FILE;
}
