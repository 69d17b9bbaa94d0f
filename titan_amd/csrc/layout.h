// Device data layout of the MI355X step engine (shared by host and kernels).
//
// Masses live in a chunked arena: a chunk is 64 consecutive mass slots, one
// wavefront's worth, and every body starts on a chunk boundary so that a
// wave never straddles two bodies.  State is FP64, double-buffered, 48 bytes
// per mass (px py pz vx vy vz, loaded as three 16-byte pairs):
//   st[(p * ncap + i) * 6 + c], p = buffer parity.
// A body's current buffer is the parity of its own step count, so bodies
// that are not stepped need no copy.  Other per-mass arrays are SoA.
//
// Springs are stored per incidence in a sliced ELLPACK order with a slice of
// one wavefront and groups of G: incidence j of mass slot `lane` in chunk c
// sits at chunk_ptr[c] + ((j / G) * 64 + lane) * G + j % G, so a lane loads
// a group's G rest lengths (and constants) as 16-byte pairs and a wave's
// loads of one group are contiguous.  The other endpoint is named by its
// arena slot (gather layout, G = 4: four uint32 per 16-byte load) or by its
// LDS row in the tile (tiled layout, G = 2: two uint16 per 4-byte load).
// Each mass lists its incident springs in ascending spring index, which
// fixes the order of the FP64 force sum.
#pragma once

#include <stdint.h>

namespace titan_mi355x {

constexpr int kChunk = 64;          // masses per chunk = wavefront size
constexpr int kWavesPerBlock = 4;   // chunks per workgroup of the gather kernel
constexpr int kBlock = kChunk * kWavesPerBlock;

// per-mass flag bits (mflags)
constexpr uint32_t kFixed = 1u;     // constraints.fixed (src/sim.cu:1100)
constexpr uint32_t kLocalLC = 2u;   // own local-constraint list
constexpr uint32_t kExt = 4u;       // extern_force may be non-zero
constexpr uint32_t kDrag = 8u;      // non-zero integer drag coefficient
constexpr uint32_t kPad = 16u;      // padding slot of a chunk (never read)

// Tiled layout (k_step_tiles): a body's masses are ordered along a Morton
// curve and cut into tiles of 1 to 4 whole chunks (up to 256 masses).  A
// tile's halo is the sorted list of masses outside it that share a spring
// with one of its masses.  The kernel stages the tile's and the halo's state
// in LDS, so a tiled body's incidences name the other endpoint by its LDS
// row: 0..nmass-1 for a tile mass, nmass + h for halo mass h.  A tile takes
// as many chunks as keep nmass + nhalo within the context's tile row budget
// (a single chunk may exceed that, up to kMaxTileRows).  Measured on the
// 1M-mass lattice, full 256-mass tiles (4 workgroups per CU) beat tiles
// capped at 682 rows (5 per CU), whose extra halo costs more than the
// occupancy gains; the default budget is therefore the maximum.
//
// A tile of an actuated body also lists its actuators: the actuated springs
// with an endpoint in the tile, in ascending spring index, each with a copy
// of its record (a, omega, phi, rest) in the tile actuator arena.  The step
// computes their actuated rest lengths at the body's clock into LDS after
// the halo rows (kActsPerRow to a row), and an actuated incidence of the
// tile names its spring's position in that list, so that the walk reads the
// value from LDS.  These rows count against the row budget like the others.
constexpr int kTileMasses = 256;
constexpr int kMaxTileRows = 2304;  // LDS budget: 2304 x 48 bytes = 108 KiB
constexpr int kTileRows = kMaxTileRows;
constexpr int kActsPerRow = 6;      // 8-byte actuated rest lengths per 48-byte LDS row
constexpr int64_t tile_act_rows(int64_t nact) { return (nact + kActsPerRow - 1) / kActsPerRow; }

struct TileDesc {
    int32_t body;
    int32_t chunk0;     // first chunk (arena index once placed)
    int32_t nmass;      // masses in the tile (<= 256)
    int32_t nhalo;
    int64_t halo_off;   // into the halo arena
    int32_t first;      // 1 for the body's first tile (advances the body clock)
    int32_t nact;       // actuators of the tile
    int64_t act_off;    // its first actuator record in the tile actuator arena
};

// Resident layout (k_rollout_resident): one workgroup runs every step of a
// stepping call for one body of at most kResidentMaxBody masses.  Its
// incidences name the other endpoint by its index in the body (uint16, bit
// 15 reserved, as the tiled layout's LDS rows, with the whole body as one
// tile and no halo).  A body of at most kResidentMax masses is held in the
// workgroup's LDS, double-buffered, when the launch has at most one body
// per compute unit; otherwise the body steps in its arena state buffers.
constexpr int kResidentMax = 1024;
constexpr int kResidentMaxBody = 32767;
constexpr int kResidentBlock = 256;
// Workgroup of the resident launch whose bodies are all in the split form:
// four threads per row of the largest split-form body (256 masses), each
// evaluating one spring term per pass (the README scene's 1,036 springs:
// one pass, and a second for 12 threads), with four waves per SIMD to hide
// the latency of a term's dependent FP64 chain.  It copies the global
// planes (at most kSplitPlanes; a launch with more uses the 256-thread
// resident kernel) into static LDS, where it also keeps each row's three
// force components; the dynamic region leaves room for both.
constexpr int kSplitMaxChunks = 4;  // a split-form body has at most this many chunks
constexpr int kSplitBlock = 1024;
constexpr int kSplitPlanes = 16;
constexpr int kSplitStaticLds = 8 * (8 * kSplitPlanes + 3 * kSplitMaxChunks * kChunk);

struct ResidentDesc {
    int32_t body;     // -1: free
    int32_t chunk0;   // first chunk (arena index)
    int32_t nmass;
    int32_t nact;     // the body's actuators: [act0, act0 + nact)
    int64_t act0;
    int64_t inc0;     // the body's incidence entries: [inc0, inc0 + ninc)
    int32_t ninc;
    int32_t cache;    // LDS bytes of a copy of its incidence data; 0: read from the arena
    int32_t scache;   // LDS bytes of its split-form region; 0: no split form
    int32_t nsp;      // split form: the body's springs (its data follows the ninc entries)
};

// A resident body whose incidence copy (or split-form region) fits this many
// bytes keeps it in LDS for all the steps of a launch.
constexpr int kResidentCacheMax = 128 * 1024;

// The split form of a resident body of at most kSplitMaxChunks chunks.  Each
// mass's incidences are in ascending spring index as its SELL entries hold
// them; the term of incidence j of row q has the slot j * nmass + q, so that
// a wave's reads of every mass's j-th term are consecutive words (a mass's
// terms kept together cost the per-mass sums many-way bank conflicts).  The
// springs are listed once each, in ascending spring index: a word with the
// rows of the endpoints a | b << 16 (b's bit 15: actuated), a word with
// their incidences' slots sa | sb << 16, then the rest lengths (or actuator
// indices), and k and damping where per spring.  This data sits in the
// body's incidence span after its SELL entries: the nsp endpoint words and
// nsp slot words in inc_nbr, the rest lengths, k and damping in inc_rest.
// A step has the workgroup's threads compute each spring's term G once,
// from endpoint a's side, and store -G in slot sa (the walk's term at a is
// G, subtracted) and +G in sb (the walk's term at b is -G exactly, zero
// signs aside); each mass's lane then adds up its own slots in order.
// LDS region: four int32 partial maxima of the degrees, the 2 nsp words,
// from a 16-byte boundary the rest lengths, k and damping where per spring,
// and the three components of the terms, (largest degree rounded up to a
// multiple of 4) x nmass slots each (8 B).
constexpr int kSplitHead = 16;
constexpr int64_t split_region_bytes(int64_t nslots, int64_t nsp, bool per_k, bool per_d) {
    return ((kSplitHead + 8 * nsp + 15) & ~(int64_t)15) + 8 * nsp * (1 + (per_k ? 1 : 0) + (per_d ? 1 : 0)) +
           24 * nslots;
}

// The body table and the arena fields a body's activation (joining the
// launches) and deactivation write, on the step stream.
struct BodyTable {
    int32_t *chunk_body;
    TileDesc *tiles;
    ResidentDesc *rdesc;
    double *bT[2];
    int64_t *bn[2], *bleft[2], *bnsync;
    int32_t *b_chunk0, *b_lc, *b_lcn;
    uint8_t *b_mode;
    double *b_k, *b_d, *b_m;
    uint8_t *b_kd;
};
struct BodyInit {
    int64_t chunk0, nchunks, tile0, ntiles, res;  // res < 0: not resident
    double t0, k, d, m;
    int32_t bid, lcb, lcn;
    uint8_t mode, kd, on;  // on: 1 joins the launches, 0 leaves them
};

// Integrators (the reference's default Euler branch and its VERLET and RK2
// build options, src/sim.cu:1141-1167).  RK2 takes two launches per step
// (stages 0 and 1, by the parity of the body's launch count); every launch
// advances the clock by hdt (dt, or dt / 2 for RK2).
constexpr int32_t kIntEuler = 0, kIntVerlet = 1, kIntRK2 = 2;

// per-body layout (b_mode)
constexpr uint8_t kModeGather = 0, kModeTiled = 1, kModeResident = 2;

// incidences per SELL group: gather (four uint32 slots in one 16-byte load)
// and tiled (two uint16 LDS rows in one 4-byte load)
constexpr int kGroupGather = 4;
constexpr int kGroupTiled = 2;

// An actuated incidence is marked by the top bit of its name (bit 31 of a
// gather slot, bit 15 of a 16-bit LDS row); its rest-length entry then holds
// an int64 in the same 8 bytes: for a tiled body, the spring's position in
// its tile's actuator list (the step reads the actuated rest length from
// LDS); otherwise the actuator's arena index, whose actuated rest length the
// step reads from the factor buffer.
constexpr uint32_t kActBit32 = 0x80000000u;
// b_kd bit 2: every mass of the body has the same m, which the tiled step
// reads from the body table instead of per mass
constexpr uint8_t kKdUniformM = 4u;
constexpr uint32_t kActBit16 = 0x8000u;

// local constraint entry (a reference CudaContactPlane / CudaBall /
// CudaConstraintPlane / CudaDirection, src/object.cu:53-141)
struct LocalC {
    double v[3];   // unit normal, ball centre or unit tangent
    double num;    // offset, radius or friction
    int32_t kind;  // TITAN_CONSTRAINT_PLANE .. TITAN_DIRECTION
    int32_t pad;
};

constexpr int64_t kMaxMassSlots = (int64_t)1 << 26;
constexpr int kBallGridMin = 8;               // balls from which a cell grid culls them
constexpr int64_t kBallGridMaxCells = 1 << 21;  // 48 * slots < 4 GiB per buffer

struct StepArgs {
    // mass arena
    double *st;               // 2 x ncap x 6
    int64_t ncap;
    const double *m;
    const uint32_t *mflags;
    const int32_t *deg;       // incident springs per mass
    const double *ext;        // 3 x ncap or null
    const int32_t *drag;      // ncap or null
    const int64_t *mlc;       // per-mass local list: begin << 20 | count
    double *acc;              // 3 x ncap
    // chunk arrays
    const int32_t *chunk_body;
    const int64_t *chunk_ptr;  // first incidence entry of the chunk
    const int32_t *chunk_ng;   // incidence groups of every lane of the chunk
    int64_t chunk_base, nchunks;  // launch range: chunks [chunk_base, nchunks)
    // incidences
    const uint32_t *inc_nbr;
    const double *inc_rest;
    const double *inc_k;      // per incidence, for bodies with per-spring k (else null)
    const double *inc_damp;   // per incidence, for bodies with per-spring damping (else null)
    const double *act;        // 4 per actuator: a, omega, phi, rest
    const double *tact;       // tiled bodies: 4 per tile actuator (the same record), by tile
    // integrator; RK2: the state a first stage leaves for the second (pos,
    // vel after the constraints: 6 per mass, as the state buffers)
    int32_t integ;
    double hdt;
    double *rk;
    // actuated rest lengths rest (1 + a sin(omega T + phi)), one per
    // actuator, for the step that reads launch slot s (ract[s]); each chunk
    // computes its share of its body's actuators, [chunk_act0, chunk_act0 +
    // chunk_nact), for the next step
    double *ract[2];
    const int64_t *chunk_act0;
    const int32_t *chunk_nact;
    // bodies: dynamic fields double-buffered by launch slot
    double *bT[2];
    int64_t *bn[2];
    int64_t *bleft[2];
    const int32_t *b_chunk0;
    // per body: spring constant and damping when every spring of the body
    // shares them, else read per incidence (b_kd bit 0: k, bit 1: damping)
    const double *b_k, *b_d;
    const uint8_t *b_kd;
    const double *b_m;        // per body: the mass of every mass of the body, when they are all equal (b_kd bit 2)
    const int32_t *b_lc;      // shared local list: begin
    const int32_t *b_lcn;     // shared local list: count
    const LocalC *lc;
    // globals
    const double *planes;     // 8 per plane: nx ny nz d mu_k mu_s 0 0
    int32_t nplanes;
    const double *balls;      // 4 per ball: cx cy cz r
    int32_t nballs;
    // uniform grid over the balls (null below kBallGridMin balls): cell
    // (ix, iy, iz) lists, in creation order, the balls that can touch it,
    // each as a copy of its record (cx cy cz r), so that a mass's test of a
    // listed ball loads it directly; one spare record ends the lists
    const int32_t *bg_start;  // nx*ny*nz + 1 offsets into bg_rec (in records)
    const double *bg_rec;
    double bg_x0, bg_y0, bg_z0, bg_inv_h;
    int32_t bg_nx, bg_ny, bg_nz;
    double gx, gy, gz, dt;
    int32_t slot;
    // tiled path
    const TileDesc *tiles;
    int64_t tile_base, ntiles;    // launch range: tiles [tile_base, ntiles)
    const int32_t *halo;      // arena mass slots
    int32_t ldim;             // LDS rows: largest nmass + nhalo + tile_act_rows(nact) of a tile
    const uint8_t *b_mode;    // per body: kModeGather, kModeTiled or kModeResident
    // resident path
    const ResidentDesc *rdesc;
    int64_t res_base, nres;   // launch range: descriptors [res_base, nres)
    int32_t res_rows;         // LDS rows per buffer: the largest resident body
    int32_t res_cache;        // LDS bytes after the two state buffers for an incidence copy
    int32_t res_split;        // 1: bodies with a split-form region use it; 2: every body does (k_rollout_split);
                              // 3: as 2, and every mass takes the lean mass pass
};

}  // namespace titan_mi355x
