/* tex2out -- text listing of a texture dump (src/tex2out.c:18-70, one rank):
 *     tex2out tex_S000100.iout tex.txt                                   */
#include <stdio.h>

#include "evp_host.h"

int main(int argc, char **argv)
{
    if (argc != 3) {
        fprintf(stderr, "usage: %s texture.iout out.txt\n", argv[0]);
        return 2;
    }
    int st = evp_texture_to_text(argv[1], argv[2]);
    if (st) fprintf(stderr, "tex2out: cannot convert %s to %s\n", argv[1], argv[2]);
    return st;
}
