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
    kStatCount = 24,
};

enum RecordKind : uint32_t { kKindDrop = 0, kKindHex = 1 };

struct DevRecord {
    uint32_t kind;
    uint32_t nt, np, qbits;
    uint64_t rays;
    uint64_t tile_rays;
    unsigned long long* bins;   // nt*np u64 fixed-point quanta
    unsigned long long* stats;  // kStatCount counters
    const double* edges;        // nt+1 u edges
    BinFrame frame;             // FP64 binning frame (e0, e1, ki)
    drop::Config drop;
    hex::Config hex;
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
