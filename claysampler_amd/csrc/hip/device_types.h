// Device-side layout of a render batch. Built by the host runtime from the
// ABI plans (include/clay_hip.h); shared by the kernels and the runtime.
//
// Layout choices for MI355X:
//  - the sample bank is int16 (half the bytes of the reference float bank);
//  - the sinc table is phase-major ([phase][tap]) so one lane reads its taps
//    as contiguous 16-byte vectors;
//  - spans are split into a hot record (mix/source kernels) and a cold record
//    (filter and pitch recurrences);
//  - control tracks are absolute point ranges in one concatenated array.
#pragma once

#include <stdint.h>

#include "../../../include/clay_hip.h"

namespace clay_dev {

constexpr int kPhases = CLAY_SINC_PHASES;
constexpr int kTailBlock = 1024;   // natural end / finalize block (src/cuda/render_types.h:92)
constexpr int kFinalizeBlocks = 32;  // tail blocks per wave of clayFinalizePcm
constexpr int kFinalizeWaves = 4;   // independent waves per block of clayFinalizePcm
constexpr int kQuantizeWaves = 4;   // windows (one per wave) per block of clayQuantizeMono
constexpr int kTileFrames = 256;   // output frames per mix tile = threads per mix block
constexpr int kStraightFrames = 1024;  // output frames per tile of a straight batch (clayMixStraight)

// Hot span record: everything the per-frame voice audio needs.
struct Span {
    uint64_t sample;                // bank offset
    double position, step, pitchSlope;
    uint64_t filtered;              // scratch offset (floats) of this span's first frame in the segment
    uint64_t phase;                 // phase scratch offset (doubles)
    int32_t begin, end, sourceEnd, loopFirst, loopEnd, loopCrossfade;
    int32_t channels, generator, stereoMatrix, variablePitch;
    uint32_t noiseSeed;
    int32_t noiseStart;
    int32_t firstMod, modCount;
    int32_t filterTrack;            // -1 = dry voice
    int32_t job;                    // batch job index
    int32_t next;                   // next span of the same voice chain, -1 at the end
    int32_t attackStart, holdStart, decayStart, releaseStart, fadeStart;
    float start, attackStep, decayRate, sustain, releaseValue, releaseRate, fadeValue, fadeStep;
    float gain, pan, left, right;
    float widthDirect, widthCross, positionLeft, positionRight, width, stereoPosition;
    // 1: a mono sampled dry voice whose audio no modulation reaches (no
    // generator, variable pitch or stereo matrix): the mix renders it with the
    // integer FIR of the packed int16 weight table.
    int32_t plain;
    // Filtered voices: frames from here on have no scratch (a bound from the
    // predicted tail of the chain; INT32_MAX = the whole span). A chain that
    // reaches it without retiring is stopped there and flagged (overflow).
    int32_t scratchEnd;
    // 1: a plain voice with no loop and no bend at positions in [0, 2^30) that
    // never fall (straight: the mix's integer FIR at exact positions).
    int32_t straight;
    int32_t pad;
};

// Cold span record: filter slots and the filter/pitch envelopes.
struct SpanCold {
    clay_voice_filter filters[CLAY_FILTER_SLOTS];
    clay_envelope_shape extra[2];
    int32_t active[2];
    int32_t sampleRate;
    int32_t pad;
};

// Modulation binding with absolute control point ranges (first < 0: constant).
struct Mod {
    int32_t first, last;
    int32_t target;
    float depth;
    int32_t shapeFirst, shapeLast;
};

struct Job {
    uint64_t mix;                   // frame offset of the job's mix in the batch mix buffer
    int32_t frames;
    int32_t controls[4];            // point offsets: CC7 [c0,c1), CC11 [c1,c2), CC10 [c2,c3)
    int32_t musicEnd, limit, fade;
    int32_t blockFirst;             // first 1024-frame block of the job
    int32_t windowFirst;            // first 50 ms window of the job
    int32_t bodyFrames, bodyFirst, onsetFirst, onsetCount;
    int32_t pad;
};

struct Tile {
    int32_t job, frame, frames, first, count;
};

// Tile claims of one mix launch (mix.h, TileClaims): a counter per XCD, each
// on a cache line of its own (device-scope atomics on one line would queue
// behind each other), cleared by clayClearQueues before the launch.
constexpr int kXcds = 8;  // MI355X: 8 XCDs, workgroups dispatched round-robin over them
struct MixQueue {
    struct alignas(128) Counter {
        uint32_t next;
        uint32_t pad[31];
    } xcd[kXcds];
};

// Filter and pitch state carried across segments (src/cuda/render_types.h:22-28).
// The recurrences run in FP64: a float32 biquad with a 200 Hz cutoff at
// 44.1 kHz (poles near z = 1) amplifies its own rounding by ~1/(1-r)^2, and a
// float pitch increment drifts the source phase over long notes; FP64 keeps
// both within the +-1 LSB bar against the FP64 transcription.
// Besides the reference fields, the last filter target of each slot and the
// inputs it was computed from are kept, so a chunk after a segment boundary
// takes exactly the path it takes without the boundary (bit-identical forms).
struct VoiceState {
    double position, step;
    int32_t pitchSpan, previousLoopEnd, previousLoopStart;
    int32_t pitchInitialized, filterInitialized;
    int32_t overflow;  // 1: the chain reached its scratch bound before it retired
    double z1[CLAY_FILTER_SLOTS][3][2], z2[CLAY_FILTER_SLOTS][3][2];
    double coefficients[CLAY_FILTER_SLOTS][5];
    double target[CLAY_FILTER_SLOTS][5];
    double input[CLAY_FILTER_SLOTS][3];
    int32_t have[CLAY_FILTER_SLOTS];
    // Voice-silence policy: retired chains end at `silence`; the chunk being
    // measured (index + 1, 0 = none) with its running peak and eligibility.
    int32_t retired, silence, silentChunks, openChunk;
    float openPeak;
    int32_t openEligible;
};

// Device pointers and the frame window of one resident unit (a batch of one
// or more jobs) for one round of render launches. Tiles of a unit are sorted
// by frame, so the tiles of a window are one contiguous range.
struct StereoStats {
    double power;
    float peak;
    int32_t pad;
};

struct UnitView {
    const Span* spans;
    const SpanCold* cold;
    const Mod* mods;
    const clay_control_point* points;
    const Job* jobs;
    const Tile* tiles;
    const int* voices;
    const int* chainHeads;
    const int* pitchHeads;
    double* phase;
    float* filtered;
    VoiceState* states;
    void* mix;
    // A straight unit (every voice straight) has 1024-frame tiles and no
    // 256-frame ones; any other unit the reverse.
    const Tile* straightTiles;
    const int* straightVoices;
    // Power and peak of each 1024-frame block, written by the straight mix
    // when this round renders the unit whole (else null: clayScanStereo).
    StereoStats* stereo;
    int16_t* mono;  // with `stereo`: the unit-scale mono PCM16 (clayQuantizeMono keeps it)
    int32_t segFirst, segEnd;           // frame window of this round
    int32_t tileFirst, tileCount;       // tiles overlapping the window
    int32_t straightFirst, straightCount;
    int32_t chainCount, pitchCount;
    // Exclusive prefix sums over the round's units (work item -> unit search).
    int32_t tileBase, straightBase, chainBase, pitchBase;
};

// Totals of one round, read by the grid-stride loops of the render kernels.
struct RoundHeader {
    int32_t units, tiles, chains, pitches;
    float silence;                      // voice-silence threshold (linear), 0 = off
    int32_t straightTiles;
    int32_t pad[2];
};


struct MonoStats {
    double power;
    int32_t peak;
    int32_t pad;
};

struct PcmJob {
    uint64_t sourceOffset;
    int32_t blockFirst;
    int32_t pad;
};

struct TailConfig {
    int32_t window, quietFrames, padding, fade;
    double relative, floor, ceiling;
};

struct AudioJobDesc {
    int32_t bodyFrames, bodyFirst, onsetFirst, onsetCount;
};

} // namespace clay_dev
