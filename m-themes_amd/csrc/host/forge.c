/* forge.c -- the reference's executable (src/evp.c:7-32) on the MI355X loop:
 * `forge input.in`.  EVP_PRECISION=mixed selects the FP32-FFT / FP64-Newton
 * loop with FP64 correction sweeps; the default is all-FP64.  EVP_LAW=dd
 * runs the dislocation-density law the reference's build compiles in; the
 * default is the power law of the configuration files.                    */
#include <assert.h>
#include <stdlib.h>
#include <string.h>

#include "evp_host.h"

int main(int argc, char **argv)
{
    assert(argc > 1);
    const char *p = getenv("EVP_PRECISION");
    if (p && strcmp(p, "mixed") == 0) evp_set_precision(EVPGPU_MIXED);
    const char *l = getenv("EVP_LAW");
    if (l && strcmp(l, "dd") == 0) evp_set_law(EVP_LAW_DD);
    GetNameList(argv);
    PrintNameList();
    OpenFiles();
    SetupJob();
    Evolution();
    DestroyJob();
    CloseFiles();
    return evp_status();
}
