/**
 * @file morefit.hh
 * Umbrella header (reference morefit.hh:20-31): data model, graph, densities, compute
 * boundary with the MI355X and CPU backends, fitter, generator, toy studies, batched fits
 * and the fit service.  The HIP backend needs libmorefit_hip.so at link time.
 */
#ifndef MOREFIT_HH
#define MOREFIT_HH

#include "morefit/dimensionvector.hh"
#include "morefit/parametervector.hh"
#include "morefit/eventvector.hh"
#include "morefit/graph.hh"
#include "morefit/pdf.hh"
#include "morefit/physicspdfs.hh"
#include "morefit/random.hh"
#include "morefit/utils.hh"
#include "morefit/compute.hh"
#include "morefit/compute_cpu.hh"
#include "morefit/compute_hip.hh"
#include "morefit/minimizer.hh"
#include "morefit/fitter.hh"
#include "morefit/generator.hh"
#include "morefit/toystudy.hh"
#include "morefit/plotter.hh"

#endif
