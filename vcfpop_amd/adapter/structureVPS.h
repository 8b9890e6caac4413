/*
 * structureVPS.h — the one declaration vcfpop's MCMC needs to hand a whole
 * run to the MI355X engine (the rest of the adapter replaces symbols vcfpop
 * already declares, src/structureCUDA.h:6-52 and src/structure.h:259-306).
 *
 * BAYESIAN<REAL>::MCMC (src/structure.cpp:2112-2166) calls it once the run
 * holds a device (ManageCUDA set usecuda):
 *
 *     ManageCUDA(false);
 *     if (usecuda) { VpsMCMC(*this); ManageCUDA(true); return; }
 *
 * It runs sweeps m = 0 .. nburnin + nreps - 1 and Arrange on the device and
 * leaves Q, MiSum (normalised), FreqM, rout, the lnL trace, nr and nlnL in the
 * run's host fields, as the reference's loop and Arrange do; WriteStructure
 * then runs unchanged.
 */
#pragma once

template <typename REAL>
struct BAYESIAN;

template <typename REAL>
void VpsMCMC(BAYESIAN<REAL>& b);
