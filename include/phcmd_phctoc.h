/* The batched-tracker entry of PHCpack's PHCtoC layer, with the reference's
 * name and signature, in libphcmd_phctoc.so (on top of libphcmd.so).
 *
 * gpumanypaths_dd (Ada/PHCtoC/gpumanypaths_dd.cpp:18-57,
 * gpumanypaths_dd.h) tracks the paths of the artificial-parameter homotopy
 * gamma (1 - t) q + t p in double-double precision: it copies the target
 * system p and then the start system q to the systems container, reads
 * their terms, reads the start solutions from the solutions container,
 * tracks every path, and replaces the solutions in the container by the
 * endpoints.  It calls the same container functions of PHCpack's C
 * interface as the reference (C/phcpack.h: copy_dobldobl_target_system_to_container,
 * copy_dobldobl_start_system_to_container; C/syscon.h:
 * syscon_number_of_symbols, syscon_number_of_dobldobl_polynomials,
 * syscon_number_of_dobldobl_terms, syscon_retrieve_dobldobl_term;
 * C/solcon.h: solcon_number_of_dobldobl_solutions,
 * solcon_retrieve_next_dobldobl_initialize,
 * solcon_retrieve_next_dobldobl_solution, solcon_clear_dobldobl_solutions,
 * solcon_append_dobldobl_solution), which the program that loads the
 * library provides (PHCpack's library, when the Ada side binds this entry
 * with pragma import as it binds the reference's).
 *
 * mode: 0 (the reference: CPU and GPU, the GPU's endpoints kept) and 2
 * (GPU) track on the GPU; 1 (CPU only) is refused with return value 1,
 * as this library has no CPU tracker.  verbose > 0 writes the sizes and a
 * summary of the path statuses to stdout.
 * The paths are tracked by phc_track with K = 2 (double-double), series of
 * degree 8, the [5/1] Pade predictor and the step-control defaults of the
 * reference's Ada tracker (phc_track_params in phcmd.h), from the program
 * that phc_program_compile builds (double-double coefficients kept as
 * high and low terms).  Each endpoint is appended with multiplicity 1,
 * t = the parameter the path reached (1 unless the path stopped), err =
 * the last Newton correction, rco = res = 0.
 * Returns 0 on success, 1 for mode 1, 2 when the containers hold no
 * square system or their sizes disagree, 3 when a library call fails
 * (the message on stderr). */
#ifndef PHCMD_PHCTOC_H
#define PHCMD_PHCTOC_H

#ifdef __cplusplus
extern "C" {
#endif

int gpumanypaths_dd(int mode, int verbose, double regamma, double imgamma);

#ifdef __cplusplus
}
#endif

#endif
