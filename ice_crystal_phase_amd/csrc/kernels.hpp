// Device-side record and tile descriptors shared by the kernels and the host
// runtime. Plain structs, no HIP runtime types.
#pragma once

#include "drop_device.hpp"
#include "hex_device.hpp"

namespace icephase {

// Per-record u64 counters written by the trace kernel.
enum StatSlot : uint32_t {
    kStatForward = 0,
    kStatBackward = 1,
    kStatSamples = 2,
    kStatInvalid = 3,
    kStatSnapped = 4,
    kStatStatus0 = 5,  // 9 slots: trace statuses 0..7, 8 = ray missed the particle
    kStatEscaped = 14,
    kStatUnresolved = 15,
    kStatInternalHits = 16,
    kStatCellAtomics = 20,   // u64 atomicAdds issued into histogram cells
    kStatDiffraction = 17,   // quanta of the Fraunhofer lobe
    kStatRayBalance = 18,    // area-rule records: max |escaped + unresolved - 1| of a ray (FP64 bits)
    kStatFaceBalance = 19,   // area-rule records: max |R + T - incident| of an interface (FP64 bits)
    kStatScratch = 23,       // scratch word of host-requested reductions (probability export)
    kStatCount = 24,
};

// Fraunhofer diffraction lobe of a hex record's projected aperture. The
// aperture is the convex polygon (counter-clockwise, up to 8 vertices) the
// prism projects onto the plane normal to the incident direction, in the
// in-plane basis (p1, p2) expressed in the binning frame's coordinates.
struct Lobe {
    double vx[8], vy[8];  // polygon vertices, mm
    d3 p1, p2;            // aperture basis (world coordinates of the binning frame)
    double k0;            // 2 pi / lambda, 1/mm
    double inv_lambda2;   // 1 / lambda^2, 1/mm^2
    double area;          // polygon area, mm^2
    double dmax;          // largest vertex distance, mm (sub-panel sizing)
    uint32_t nv, rows;    // vertex count; leading theta rows computed
    double* mass;         // rows*np FP64 cell masses (mm^2 before scaling)
    double* partial;      // per-workgroup partial sums of mass, then total and captured fraction
};

enum RecordKind : uint32_t { kKindDrop = 0, kKindHex = 1 };

// One deterministic entry of a hex record traced from an area quadrature
// instead of the counter RNG: entry point (circumradius units) on a lit face,
// and the ray's weight relative to the record's largest (qf <= 1).
struct AreaRay {
    d3 pos;
    double qf;
    uint32_t face, pad;
};

struct DevRecord {
    uint32_t kind;
    uint32_t nt, np, qbits;
    uint64_t rays;
    uint64_t tile_rays;
    unsigned long long* bins;   // nt*np u64 fixed-point quanta
    unsigned long long* stats;  // kStatCount counters
    const double* edges;        // nt+1 u edges
    BinFrame frame;             // FP64 binning frame (e0, e1, ki)
    // Quanta of an emission: rint(power * qfactor * 2^qbits). qfactor = 1 for a
    // record that owns its histogram alone; orientation nodes sharing one
    // histogram carry their sample weight relative to the largest node's.
    double qfactor;
    double rot[9];              // hex: body -> ensemble rotation (row-major) when rotated
    uint32_t rotated, lobe_on;
    uint32_t owns;              // 1: this record's histogram (zeroed with it); 0: a node adding into another's
    const AreaRay* area;        // hex, area kernel only: the entries of an area quadrature (rays of them)
    drop::Config drop;
    hex::Config hex;
    Lobe lobe;
};

// Tiles are implicit: the records of one kind are listed in `recs_of_kind`,
// and tile_start[k] is the first tile of the k-th of them (tile_start[n] =
// total tiles). Tile t of record r covers rays [(t - start) * tile_rays, ...).
struct TileMap {
    const uint32_t* recs_of_kind;
    const uint64_t* tile_start;
    uint32_t n;
    uint32_t pad;
    uint64_t total;
};

// CDF descriptors (one per record).
struct CdfJob {
    uint32_t nt, np;            // histogram grid
    uint32_t ont, onp;          // stored grid (divides nt, np)
    const unsigned long long* bins;  // integer source (or null)
    const double* mass;         // FP64 source (used when bins is null)
    const unsigned long long* stats; // pole quanta for the integer source
    const double* edges;        // nt+1
    const double* cedges;       // ont+1
    double* coarse;             // ont*onp scratch when coarsening, else null
    double* phi;                // onp+1
    double* theta;              // onp*(ont+1)
    double* cols;               // onp column sums
    double* scalars;            // kCdfScalars per record
    double* partials;           // kCdfPartial * blocks
};

enum CdfScalar : uint32_t {
    kCdfLargest = 0,
    kCdfMassFine = 1,
    kCdfAxialFine = 2,
    kCdfMassStored = 3,
    kCdfAxialStored = 4,
    kCdfTotal = 5,
    kCdfL1 = 6,
    kCdfLost = 7,
    kCdfRecovered = 8,
    kCdfAxialCdf = 9,
    kCdfMaxCell = 10,
    kCdfLostCells = 11,
    kCdfMalformed = 12,
    kCdfUnderflow = 13,
    kCdfTv = 14,
    kCdfNegative = 15,
    kCdfScalars = 16,
};

}  // namespace icephase
