// SFZ instruments: the regions a job's notes select and everything a voice
// span needs from them.
//
// Supported: <control> <global> <master> <group> <region>, #define and
// #include; sampled regions (PCM16 WAV) and the *sine, *silence and *noise
// generators; key and velocity ranges, CC conditions (locc/hicc); tuning,
// bend range and bend smoothing; volume, amplitude, velocity tracking and
// curves, pan, width and position; offset, end, delay and the loop modes with
// crossfade; the amplitude envelope with velocity offsets, and the filter and
// pitch envelopes; two filters and three EQ bands; controller bindings
// (_onccN, _curveccN, _smoothccN) with the extended sources note-on
// velocity (131), release velocity (132), key (133), keys held (134) and the
// key step (140, 141); the sustain pedal settings.
// Not supported (rejected as unsupported opcodes): LFOs, flex envelopes,
// key/velocity/controller crossfades, round robin, random selection and
// randomisation, keyswitches, choke groups and note polyphony, release,
// first, legato and CC triggers, dynamic envelopes and <curve> tables.
// Opcode names, units and ranges follow sfizz 1.2.3 as the reference accepts
// them (src/sfz/instrument.cpp:759-891).
#pragma once

#include "../../../include/clay_hip.h"
#include "audio.h"

#include <array>
#include <cstdint>
#include <map>
#include <string>
#include <utility>
#include <vector>

namespace clay {

// Modulation targets, in the order of the ABI's clay_mod_target.
enum class ModTarget : int {
    amplitude, volume, pan, width, position, pitch,
    envDelay, envAttack, envHold, envDecay, envSustain, envStart, envRelease,
    offset, delay, cutoff, resonance, cutoff2, resonance2, filterVelocity, filterVelocity2,
    lfoFrequency, lfoPhase, lfoDelay, lfoFade,
    filterEnvDelay, filterEnvAttack, filterEnvHold, filterEnvDecay, filterEnvSustain, filterEnvStart, filterEnvRelease,
    pitchEnvDelay, pitchEnvAttack, pitchEnvHold, pitchEnvDecay, pitchEnvSustain, pitchEnvStart, pitchEnvRelease,
    filterEnvDepth, pitchEnvDepth,
    eq1Frequency, eq1Bandwidth, eq1Gain, eq2Frequency, eq2Bandwidth, eq2Gain, eq3Frequency, eq3Bandwidth, eq3Gain
};

// Filter kinds and generators, in the order of the ABI enums.
enum class FilterType : int { none, low1, high1, low2, high2, band2, notch2, low4, high4, band4, low6, high6, band6, peak };
enum class Generator : int { none, sine, silence, noise };

enum class LoopMode { fromFile, none, oneShot, continuous, sustain };

// Envelope stages in this order: delay, attack, hold, decay (s), sustain,
// start (0..1), release (s).
enum EnvelopeStage { stageDelay, stageAttack, stageHold, stageDecay, stageSustain, stageStart, stageRelease, stageCount };
using Stages = std::array<float, stageCount>;

// A controller bound to a parameter: value = depth x curve(source).
struct Binding {
    ModTarget target;
    int source;      // CC 0..127 or an extended source >= 131
    float depth = 0;
    int curve = 0;   // built-in curve 0..6
    int smoothMs = 0;
};

// Filter (fileg_) or pitch (pitcheg_) envelope of a region.
struct ModEnvelope {
    bool used = false;
    Stages stages{0, 0, 0, 0, 1, 0, 0};
    Stages velocity{};  // added at velocity 127, scaled linearly
    float depth = 0, velocityDepth = 0;
};

struct FilterSettings {
    FilterType type = FilterType::none;
    float cutoff = 0, resonance = 0, keyTrack = 0, velocityTrack = 0;
    int keyCenter = 60;
};

struct EqSettings {
    bool used = false;
    float frequency = 0, bandwidth = 1, gain = 0;
};

struct CcWindow {
    int cc, low, high;
};

struct Region {
    // Source.
    int sample = -1;  // bank id; -1 for generators
    Generator generator = Generator::none;
    int offset = 0, end = -1, loopStart = -1, loopEnd = -1;
    LoopMode loop = LoopMode::fromFile;
    float loopCrossfade = .001f, oscillatorPhase = 0;
    // Selection.
    int keyLow = 0, keyHigh = 127, keyRoot = 60, velocityLow = 0, velocityHigh = 127;
    bool onNotes = true;  // false: key or hikey -1, never started by a note
    std::vector<CcWindow> conditions;
    // Pitch.
    int transpose = 0;
    float tune = 0, keyTrack = 100, bendUp = 200, bendDown = -200;
    int bendSmoothMs = 0;
    // Level and placement.
    float volumeDb = 0, amplitude = 1, velocityTrack = 1, pan = 0, width = 1, position = 0;
    std::vector<std::pair<int, float>> velocityCurve;  // sorted, ends at 0 and 127 when used
    float delay = 0;
    // Envelopes.
    Stages amp{0, 0, 0, 0, 1, 0, .001f};
    Stages ampVelocity{};
    std::array<ModEnvelope, 2> modEnvelopes;  // filter, pitch
    std::array<FilterSettings, 2> filters;
    std::array<EqSettings, 3> eq;
    std::vector<Binding> bindings;
    // Sustain pedal.
    bool sustainSwitch = true;
    int sustainCC = 64, sustainLow = 1;
};

struct Instrument {
    std::vector<Region> regions;
    std::array<std::vector<int>, 128> regionsOfKey;
    std::array<float, 128> ccDefault{};
    std::array<bool, 128> ccUsed{};
    std::map<int, float> extendedDefaults;  // set_cc for sources >= 128
};

Instrument loadInstrument(const std::string& path, AudioBank& bank);

// Built-in controller curves 0..6 of sfizz (src/sfz/instrument.cpp:909-931)
// at a controller value 0..127, interpolated between integer values.
float curveValue(int curve, float value);

} // namespace clay
