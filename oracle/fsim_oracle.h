/*
 * fsim_oracle.h — FP64 CPU oracle of the FLIP/APIC step (TEST INFRASTRUCTURE).
 *
 * This library is the checker, never the product: only tests/, the smoke
 * entry point and bench.py's CPU-baseline leg may load it.  It restates the
 * reference's algorithm (dkiovics/fluid_simulator, src/Simulator/simulator/)
 * with the reference's data structures: AoS FP64 particles and cells,
 * atomic double scatter in P2G, OpenMP loops and a serial MIC(0).
 *
 * It shares only the POD parameter/obstacle structs and enums of
 * include/fsim.h with the product, so that both are driven by the same
 * inputs.  Per-function citations are in fsim_oracle.c.
 */
#ifndef FSIM_ORACLE_H
#define FSIM_ORACLE_H

#include "../include/fsim.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fo_sim fo_sim;

fo_sim* fo_create(const fsim_grid_desc* grid, const fsim_params* params);
void fo_destroy(fo_sim* s);
/* 0 or 1 = serial loops (the reference's _DEBUG build); >1 = OpenMP threads */
void fo_set_threads(fo_sim* s, int nthreads);
void fo_set_params(fo_sim* s, const fsim_params* params);
void fo_set_particle_radius(fo_sim* s, double r);

int fo_set_particles(fo_sim* s, int64_t n, const double* pos, const double* vel, const double* c);
int64_t fo_num_particles(const fo_sim* s);
void fo_get_particles(const fo_sim* s, double* pos, double* vel, double* c);

void fo_set_obstacles(fo_sim* s, int n, const fsim_obstacle* obs);
int fo_get_obstacles(const fo_sim* s, fsim_obstacle* out);

void fo_step(fo_sim* s, double dt, fsim_step_stats* stats);
int fo_run_stage(fo_sim* s, int stage, double dt);
void fo_get_stats(const fo_sim* s, fsim_step_stats* out);

/* Float fields are exchanged as double; LABEL as uint8; counts as int32. */
int fo_get_grid(const fo_sim* s, int field, void* dst);
int fo_set_grid(fo_sim* s, int field, const void* src);
int fo_solver_apply(fo_sim* s, int op, double dt, const double* in, double* out);

/* Serial stable counting sort of the particles by grid cell:
 * count[Nc], start[Nc+1] (exclusive offsets), perm[Np] = particle ids in
 * (cell, id) order. */
void fo_bin(const fo_sim* s, int32_t* count, int32_t* start, int64_t* perm);

/* The renderer's snapshot (simulationManager.cpp:223-231): out[5*i ..] =
 * pos[3], |v| and the FP32-accumulated density of particle i. */
void fo_gfx_snapshot(const fo_sim* s, double* out);

/* SplitMix64 (Steele, Lea, Flood 2014): n-th output of the generator seeded
 * with `seed`. */
uint64_t fo_splitmix64(uint64_t seed, uint64_t n);

#ifdef __cplusplus
}
#endif
#endif
