/*
 * elec_oracle.c -- CPU restatement of the polarizable electrostatics (placeholder
 * until the Ewald / Thole restatement lands).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"
