/** \file common.hh
 * \brief Constants and fatal-error handling of the drop-in PGMG API
 * (reference: common.hh:12-65, common.cc:52-55).
 *
 * Header-only; the numerical work is done by libpgmg_amd.so through the C ABI
 * of pgmg_capi.h.  The library uses no MPI: one process drives one GPU, so the
 * reference's MPI tags have no counterpart.  A caller that includes mpi.h
 * before these headers (as the reference's drivers do) gets the reference's
 * communicator constant `world`. */
#ifndef PGMG_AMD_COMMON_HH
#define PGMG_AMD_COMMON_HH

#include <cstdio>
#include <cstdlib>
#include <limits>

#include "../pgmg_capi.h"

#ifdef MPI_VERSION
/** The communicator of the reference's drivers (common.hh:17), e.g. for
 * MPI_Barrier(world) in mg_parallel.cc:43.  No library call takes it. */
const MPI_Comm world = MPI_COMM_WORLD;
#endif

/** The maximum number of multigrid levels (common.hh:15). */
const int mg_max_levels = 16;

/** Error codes of the reference (common.hh:19-26); the C ABI returns each
 * plus one (PGMG_ERR_FILE == PGMG_FILE_ERROR + 1, ...). */
enum {
	PGMG_FILE_ERROR,
	PGMG_MEMORY_ERROR,
	PGMG_INTERNAL_ERROR,
	PGMG_MULTIGRID_ERROR,
	PGMG_MATH_ERROR,
	PGMG_SETUP_ERROR
};

/** Linear system tolerance of the bottom LU test (common.hh:58). */
const double pgmg_exact_tol = std::numeric_limits<double>::epsilon() * 1e4;
/** Gauss-Seidel sweeps of the bottom solve without LU (common.hh:62). */
const int pgmg_gs_exact = 24;
/** Largest bottom level eligible for the LU (common.hh:65). */
const int pgmg_max_exact = 2048;

/** Prints the message and ends the process with the status, as the
 * reference's MPI_Abort does (common.cc:52-55). */
inline void p_fatal_error(const char *p, int status) {
	fprintf(stderr, "%s\n", p);
	fflush(stderr);
	std::exit(status);
}

/** Same, for non-MPI helpers (common.cc:42-45). */
inline void fatal_error(const char *p, int code) {
	fprintf(stderr, "Error: %s\n", p);
	std::exit(code);
}

inline void p_die() {p_fatal_error("Died", PGMG_INTERNAL_ERROR);}

/** A failed C ABI call: report the library's message and end the process
 * with the reference's code for that class of error. */
inline void pgmg_check(int rc, const char *what) {
	if (rc == PGMG_OK) return;
	char buf[640];
	snprintf(buf, sizeof buf, "%s: %s", what, pgmg_last_error());
	p_fatal_error(buf, rc >= 1 && rc <= 6 ? rc - 1 : PGMG_INTERNAL_ERROR);
}

#endif
