// Kernel argument block shared by the host library and the JIT-compiled solver kernels.
// Passed by value; every field has a fixed size so host (x86-64) and device (amdgcn)
// layouts agree.
#pragma once

struct deg_kargs {
    long long n_traj;         // trajectories in this launch
    long long stride;         // leading dimension of every SoA array (>= id_offset + n_traj)
    long long id_offset;      // first trajectory id of this launch (chunked launches)
    const float* u0;          // [n][stride]
    const float* p;           // [np][stride]
    const float* t0s;         // [stride] or null -> t0
    const float* tfs;         // [stride] or null -> tf
    const float* saveat;      // [nsaveat] or null
    float* ts;                // [nsave][stride]
    float* us;                // [nsave][n][stride]
    int* retcode;             // [stride] or null
    int* naccept;             // [stride] or null
    int* nreject;             // [stride] or null
    unsigned int* queue;      // persistent work counter (one 32-bit word), zeroed before launch
    const void* mlp;          // learned step proposer weights (packed), or null
    float* steplog;           // attempt log rows for training, or null
    unsigned int* steplog_count;  // atomic row counter for steplog
    long long max_attempts;   // per-trajectory attempt cap
    long long steplog_cap;    // rows available in steplog
    float t0, tf, dt, abstol, reltol;
    int nsaveat;
    int nsave;
    int save_everystep;
    int refill;               // persistent grids: idle lanes per wave that trigger a refill (1..64)
    int chunk;                // persistent grids: ids claimed per queue atomic
    int mlp_every;            // learned: proposer evaluated every k-th loop iteration
    const int* order;         // optional start order: k-th trajectory started = id_offset + order[k]
    const float* tstops;      // [ntstops] stopping times shared by all trajectories, or null
    int ntstops;
    int pad0;
    const unsigned long long* seeds;  // SDE: [stride] per-trajectory seeds, or null -> seed
    unsigned long long seed;          // SDE: shared seed
    long long traj_base;              // SDE: ensemble index of column 0 (the noise stream id is traj_base + column)
    const float* inputs_ordered;      // optional with order: (n + np) columns of 8 ceil(n_traj / 8), queue-major
};
