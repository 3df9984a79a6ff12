/*
 * music_gcmc_oracle.c -- CPU FP64 restatement of the GCMC move loop
 * (test infrastructure only; see music_oracle.h).
 */
#include "music_oracle.h"
