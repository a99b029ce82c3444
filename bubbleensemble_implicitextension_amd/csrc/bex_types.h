// Shared host/device definitions of the gfx950 per-block ensemble solver.
#pragma once

#include <stdint.h>

struct BexMlpDevice;

// Per-system counters (same layout as the CPU oracle's).
enum BexStat {
  BEX_STAT_ACCEPTED = 0,
  BEX_STAT_REJECTED = 1,
  BEX_STAT_STAGES = 2,
  BEX_STAT_BICG_ITERATIONS = 3,
  BEX_STAT_BICG_RHO_EXITS = 4,
  BEX_STAT_BICG_MAXIT = 5,
  BEX_STAT_EVENTS = 6,
  BEX_STAT_FORCED_ACCEPTS = 7,
  BEX_STAT_TERMINATED = 8,
  BEX_STAT_EVENTS_OMITTED = 9,
  BEX_STAT_ATTEMPTS = 10,
  BEX_STAT_HIST0 = 11,  // 20 bins: stages that ran k = 0..19 BiCG iterations (19 = 19 or more)
  BEX_STAT_WARM_FALLBACKS = 31,  // learned initial guess rejected or not converged: restarted
  BEX_NSTAT = 32
};

// Everything one launch of the solver kernel reads: buffer pointers in the reference's
// structure-of-arrays layout (unit scope [k][NS*N], system scope [k][NS]) and the options.
struct BexKernelArgs {
  int NS;
  int ActiveSystems;
  int NDO;
  int first_system;  // systems [first_system, first_system + grid*SPB) are handled
  long long TLT;     // total logical threads of the reference layout
  int ref_spb;       // systems per block of the reference layout (Interface.cuh:875-879)
  long long ref_block_stride;  // logical threads per block incl. padding

  double* TimeDomain;
  double* ActualState;
  double* ActualTime;
  double* UnitParameters;
  double* SystemParameters;
  double* GlobalParameters;
  int* IntegerGlobalParameters;
  double* UnitAccessories;
  int* IntegerUnitAccessories;
  double* SystemAccessories;
  int* IntegerSystemAccessories;
  const double* CouplingMatrix;   // coupling 0, column-major [row + col*N]
  const double* CouplingMatrixT;  // its transpose, column-major
  int* DenseOutputIndex;
  double* DenseOutputTimeInstances;
  double* DenseOutputStates;
  const double* RelativeTolerance;
  const double* AbsoluteTolerance;
  const double* EventTolerance;
  const int* EventDirection;
  long long* Stats;  // [NS][BEX_NSTAT], accumulated
  long long* Timeline;  // optional [NS][2]: device wall-clock ticks at system start and end
  // Host-mapped stop flag: while it is nonzero, every running system ends at its next check
  // (every 256 step attempts) with its state and time written back as they stand.
  const int* StopFlag;
  // > 0: a system ends at its first check (every 256 step attempts) after this many attempts in
  // the launch, state and time written back as they stand.
  long long MaxStepAttempts;

  double InitialTimeStep;
  double MaximumTimeStep;
  double MinimumTimeStep;
  double TimeStepGrowLimit;
  double TimeStepShrinkLimit;
  double DenseOutputMinimumTimeStep;
  int DenseOutputSaveFrequency;
  double BiCGRhoTolerance;
  double BiCGResidualTolerance;
  int BiCGMaxIterations;
  int PriorityStepThreshold;  // accepted steps after which a system's wave priority rises

  // BiCG warm start and training-sample collection (no reference counterpart).
  const BexMlpDevice* WarmStart;     // network in the solver's operand layout (learned kernel)
  const unsigned char* WarmStartSystems;  // [NS] learned kernel: 0 = plain start for that
                                          // system; nullptr: every system takes the learned start
  float* Samples;                    // [SampleCapacity][BEX_SAMPLE_FLOATS]
  unsigned long long* SampleCount;   // rows reserved so far (may exceed the capacity)
  long long SampleCapacity;
  int SampleEvery;                   // a system records every SampleEvery-th RK stage

  // Interpolated dense output (no reference counterpart).
  int DenseInterpolation;  // 0: step-point sampling (reference), 1: Hermite at a uniform grid
  double* DenseAux;        // [2 UD][TLT]: start state and derivative of the pending step
};
