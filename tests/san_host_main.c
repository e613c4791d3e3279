/* Sanitizer driver (test infrastructure): the oracle on a small job under
 * both laws, and the drop-in's host parsers on the same input, built with
 * -fsanitize=address,undefined by tests/test_sanitize_cpu.py. */
#include <stdio.h>
#include <string.h>

#include "../oracle/evp_oracle.h"
#include "../m-themes_amd/csrc/host/evp_host.h"

int main(int argc, char **argv)
{
    if (argc < 3) return 2;
    for (int law = 0; law < 2; law++) {
        orc_set_law(law);
        int st = 0;
        orc_job *J = orc_job_setup(argv[1], &st);
        if (!J) return 10 + st;
        int it[2], forced[2] = {20, 20};
        st = orc_job_run(J, 2, it, forced);
        orc_job_free(J);
        if (st) return 20 + st;
    }
    evp_namelist nl;
    memset(&nl, 0, sizeof(nl));
    if (evp_parse_namelist(argv[1], &nl) < 0) return 30;
    evpgpu_phase ph;
    memset(&ph, 0, sizeof(ph));
    if (evp_parse_sx(argv[2], nl.Slip_PhaseI, &ph)) return 31;
    evp_dd_keys dd;
    if (evp_parse_dd_keys(argv[1], &dd)) return 32;
    evpgpu_dd_phase out;
    evp_dd_phase_setup(&dd, 0, &ph, 1.0e5, 0.35, &out);
    printf("ok %d %g\n", ph.nsys, out.chi[0][0]);
    return 0;
}
