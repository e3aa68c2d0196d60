// CPU planner: MIDI events + SFZ regions -> voice spans and control tracks
// (the RenderPlan of src/render/plan.h:9-23 in the reference).
#pragma once

#include "../../../include/clay_hip.h"
#include "instrument.h"
#include "midi.h"

#include <vector>

namespace clay {

struct Plan {
    std::vector<clay_voice_span> spans;
    std::vector<std::vector<clay_control_point>> controls;  // tracks 0,1,2 = CC7, CC11, CC10
    std::vector<clay_modulation> modulation;                // plan-local track indices
    int filterTracks = 0;
    int frames = 0;
    int musicEnd = 0, limit = 0, fadeFrames = 0;
    bool tailClipped = false;
    int bodyFrames = 0;
    std::vector<int> onsets;
    uint64_t volumeGroup = 0;
};

// src/render/plan.cpp:1160-1164
Plan planRender(const Instrument& instrument, AudioBank& bank, const Midi& midi, int rate, unsigned seed, int limit,
                TailPolicy tail);

// Flat view for the C ABI; `storage` keeps the concatenated points alive.
struct PlanView {
    clay_plan_view view{};
    std::vector<clay_control_point> points;
    std::vector<int32_t> offsets;
};
void makeView(const Plan& plan, PlanView& out);

} // namespace clay
