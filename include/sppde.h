/* Forwarding header: SW2's driver includes "sppde.h" (main/Main.c:18); the
 * sppde helpers it calls (pprintf, CRASH, cr_*, quants, checkr) are declared
 * in sw2_hip.h. */
#include "sw2_hip.h"
