/* Forwarding header: SW2's driver includes "sw.h" (main/Main.c:20); the
 * world struct and the initialize/loop/destroy entry points it names are
 * declared in sw2_hip.h. */
#include "sw2_hip.h"
