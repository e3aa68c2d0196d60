// C ABI of the host planner (include/clay_host.h).
#include "../../../include/clay_host.h"
#include "audio.h"
#include "instrument.h"
#include "midi.h"
#include "plan.h"

#include <climits>
#include <cmath>
#include <cstdio>
#include <deque>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>

struct clay_session {
    clay_session_config cfg{};
    clay::AudioBank bank;
    std::map<std::string, clay::Instrument> instruments;
    std::deque<clay::Plan> plans;
    std::deque<clay::PlanView> views;
    std::string error;
    std::string lastMidiPath;
    clay::Midi lastMidi;
};

namespace {

clay::TailPolicy policyOf(int p)
{
    return p == CLAY_POLICY_ORIG ? clay::TailPolicy::orig : p == CLAY_POLICY_CUSTOM ? clay::TailPolicy::custom
                                                                                      : clay::TailPolicy::natural;
}

} // namespace

extern "C" {

int clay_session_create(const clay_session_config* config, clay_session** out)
{
    if (!config || !out || config->rate < 8000 || config->rate > 192000)
        return CLAY_E_INVALID;
    auto s = new clay_session;
    s->cfg = *config;
    *out = s;
    return CLAY_OK;
}

int clay_session_destroy(clay_session* s)
{
    delete s;
    return CLAY_OK;
}

// Planning of one job as in src/main.cpp:323-353.
int clay_session_add_job(clay_session* s, const char* sfz_path, const char* midi_path, uint64_t group, size_t* index)
{
    if (!s || !sfz_path || !midi_path)
        return CLAY_E_INVALID;
    try {
        const auto& cfg = s->cfg;
        const auto tail = policyOf(cfg.tail_policy);
        auto inst = s->instruments.find(sfz_path);
        if (inst == s->instruments.end())
            inst = s->instruments.emplace(sfz_path, clay::loadInstrument(sfz_path, s->bank)).first;
        if (s->lastMidiPath != midi_path) {
            s->lastMidi = clay::loadMidi(midi_path, cfg.rate, tail, cfg.measure != 0);
            s->lastMidiPath = midi_path;
        }
        const auto& midi = s->lastMidi;
        const int musicEnd = midi.end + 1;
        int limit = INT_MAX;
        if (tail == clay::TailPolicy::natural) {
            limit = musicEnd + 120 * cfg.rate;
            if (cfg.max_audio_seconds >= 0)
                limit = std::min(limit, int(std::floor(cfg.max_audio_seconds * cfg.rate)));
            if (limit < musicEnd)
                throw std::runtime_error("natural render exceeds the audio duration limit");
        } else if (cfg.tail_max_seconds >= 0) {
            limit = musicEnd + int(std::floor(cfg.tail_max_seconds * cfg.rate));
        }
        auto plan = clay::planRender(inst->second, s->bank, midi, cfg.rate, 1, limit, tail);
        plan.musicEnd = musicEnd;
        plan.limit = limit;
        const int customFade = int(std::lround((cfg.tail_fade_ms < 0 ? 100 : cfg.tail_fade_ms) * (cfg.rate / 1000.)));
        if (tail == clay::TailPolicy::custom && plan.tailClipped && plan.frames == limit)
            plan.fadeFrames = std::min(customFade, std::max(0, plan.frames - musicEnd));
        plan.volumeGroup = group;
        plan.bodyFrames = midi.bodyFrames;
        if (cfg.keep_onsets)
            plan.onsets = midi.onsets;
        s->plans.push_back(std::move(plan));
        s->views.emplace_back();
        clay::makeView(s->plans.back(), s->views.back());
        if (index)
            *index = s->plans.size() - 1;
        return CLAY_OK;
    } catch (const std::exception& e) {
        s->error = e.what();
        return CLAY_E_INVALID;
    }
}

const clay_plan_view* clay_session_plan(clay_session* s, size_t index)
{
    if (!s || index >= s->views.size())
        return nullptr;
    return &s->views[index].view;
}

size_t clay_session_job_count(const clay_session* s) { return s ? s->plans.size() : 0; }

const int16_t* clay_session_bank(const clay_session* s, uint64_t* count)
{
    if (!s)
        return nullptr;
    if (count)
        *count = s->bank.pcm.size();
    return s->bank.pcm.data();
}

int clay_session_reset(clay_session* s)
{
    if (!s)
        return CLAY_E_INVALID;
    s->plans.clear();
    s->views.clear();
    s->bank.releaseDecoded();
    s->lastMidiPath.clear();
    return CLAY_OK;
}

const char* clay_session_last_error(const clay_session* s) { return s ? s->error.c_str() : ""; }

int clay_midi_load(const char* path, int rate, int tail, int measure, int* events, int capacity, int* count, int* end,
                   int* bodyFrames, int* onsets, int onsetCapacity, int* onsetCount, char* error, int errorCapacity)
{
    try {
        const auto midi = clay::loadMidi(path, rate, policyOf(tail), measure != 0);
        *count = int(midi.events.size());
        *end = midi.end;
        *bodyFrames = midi.bodyFrames;
        *onsetCount = int(midi.onsets.size());
        for (int i = 0; i < *count && i < capacity; ++i) {
            const auto& e = midi.events[size_t(i)];
            events[4 * i] = e.frame;
            events[4 * i + 1] = int(e.kind);
            events[4 * i + 2] = e.a;
            events[4 * i + 3] = e.b;
        }
        for (int i = 0; i < *onsetCount && i < onsetCapacity; ++i)
            onsets[i] = midi.onsets[size_t(i)];
        return CLAY_OK;
    } catch (const std::exception& e) {
        std::snprintf(error, size_t(errorCapacity), "%s", e.what());
        return CLAY_E_INVALID;
    }
}

size_t clay_abi_sizeof(const char* name)
{
    const std::string n = name ? name : "";
#define CLAY_SIZE(T) if (n == #T) return sizeof(T);
    CLAY_SIZE(clay_voice_filter) CLAY_SIZE(clay_envelope_shape) CLAY_SIZE(clay_voice_span)
    CLAY_SIZE(clay_control_point) CLAY_SIZE(clay_modulation) CLAY_SIZE(clay_plan_view)
    CLAY_SIZE(clay_audio_config) CLAY_SIZE(clay_audio_metrics) CLAY_SIZE(clay_tail_result)
    CLAY_SIZE(clay_hip_config) CLAY_SIZE(clay_render_batch) CLAY_SIZE(clay_hip_stats)
    CLAY_SIZE(clay_session_config)
#undef CLAY_SIZE
    return 0;
}

int clay_write_wave(const char* path, const int16_t* pcm, int32_t data_frames, int32_t frames, int32_t rate,
                    int32_t channels)
{
    try {
        clay::writeWave(path, pcm, data_frames, frames, rate, channels);
        return CLAY_OK;
    } catch (const std::exception&) {
        return CLAY_E_INVALID;
    }
}

int clay_write_flac(const char* path, const int16_t* pcm, int32_t data_frames, int32_t frames, int32_t rate,
                    int32_t channels, int32_t level)
{
    try {
        clay::writeFlac(path, pcm, data_frames, frames, rate, channels, level);
        return CLAY_OK;
    } catch (const std::exception&) {
        return CLAY_E_INVALID;
    }
}

int clay_read_flac(const char* path, int32_t* rate, int32_t* channels, int16_t* pcm, uint64_t capacity, uint64_t* samples)
{
    int r = 0, c = 0;
    std::vector<int16_t> data;
    if (!clay::readFlac(path, r, c, data))
        return CLAY_E_INVALID;
    if (rate)
        *rate = r;
    if (channels)
        *channels = c;
    if (samples)
        *samples = data.size();
    if (pcm) {
        if (capacity < data.size())
            return CLAY_E_INVALID;
        std::copy(data.begin(), data.end(), pcm);
    }
    return CLAY_OK;
}

} // extern "C"
