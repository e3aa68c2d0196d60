// CPU planner. It walks the frame-ordered MIDI events, keeps controller,
// pedal, keyswitch and voice state, and describes every sounding voice as one
// or more spans with fixed playback parameters plus piecewise-linear control
// tracks. Semantics (and therefore every span the renderer sees) follow
// src/render/plan.cpp of the reference; random draws use the same
// std::mt19937 streams and seeds so region selection is identical.
#include "plan.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <limits>
#include <map>
#include <random>
#include <stdexcept>
#include <tuple>

namespace clay {
namespace {

constexpr int kForever = std::numeric_limits<int>::max();
constexpr float kSilence = 1e-4f;      // release level that starts the final 50 ms fade
constexpr int kBendBlock = 1024;       // sfizz_render block used for bend/controller ramps
constexpr float kMasterGainDb = -7.35f;
constexpr float kQuarterPi = .7853981633974483f;

enum class VoiceKind { attack, release, controller };

struct Voice {
    clay_voice_span span;
    int firstLfo = 0, lfoCount = 0;
    int key = 0, velocity = 0, triggerFrame = 0;
    const Region* region = nullptr;
    double baseStep = 0;
    float releaseSeconds = 0;
    bool noteOff = false;
    VoiceKind kind = VoiceKind::attack;
    bool choked = false, sostenuto = false;
    std::vector<int> flexShapes;
};

struct DelayedRelease {
    int key, velocity, duration;
    const Region* region;
    bool sostenuto;
};

struct Bend {
    int begin = 0, end = 0, from = 0, to = 0;
};

struct LfoInput {
    int track;  // -1: constant `depth`
    ModTarget target;
    float depth;
};

struct LfoTrack {
    const Lfo* source;
    int begin, end, delay, fade;
    std::vector<LfoInput> inputs;
    std::vector<std::pair<ModTarget, int>> outputs;
};

using Track = std::vector<clay_control_point>;

float trackAt(const Track& track, int frame)
{
    auto it = std::upper_bound(track.begin(), track.end(), frame,
                               [](int f, const clay_control_point& p) { return f < p.frame; });
    const auto& p = it == track.begin() ? track.front() : *std::prev(it);
    return p.value + float(frame - p.frame) * p.step;
}

float lfoWave(int wave, double phase)
{
    const float p = float(phase - std::floor(phase));
    switch (wave) {
    case 0: return p < .25f ? 4 * p : p > .75f ? 4 * p - 4 : 2 - 4 * p;
    case 1: {
        const float x = 2 * p - 1;
        return -4 * x * (1 - std::abs(x));
    }
    case 2: return p < .75f;
    case 3: return p < .5f;
    case 4: return p < .25f;
    case 5: return p < .125f;
    case 6: return 2 * p - 1;
    case 7: return 1 - 2 * p;
    default: return 0;
    }
}

// amp_veltrack with the default (v/127)^2 or amp_velcurve_N points (plan.cpp:80-97).
float velocityGain(const Region& r, int velocity)
{
    const float normalized = float(velocity) / 127;
    float gain = normalized * normalized;
    if (!r.velocityCurve.empty()) {
        const auto hi = std::lower_bound(r.velocityCurve.begin(), r.velocityCurve.end(), velocity,
                                         [](const auto& p, int v) { return p.first < v; });
        if (hi->first == velocity) {
            gain = hi->second;
        } else {
            const auto lo = std::prev(hi);
            const float t = float(velocity - lo->first) / float(hi->first - lo->first);
            gain = lo->second + t * (hi->second - lo->second);
        }
    }
    const float track = std::abs(r.velocityTrack) * (1 - gain);
    return r.velocityTrack < 0 ? track : 1 - track;
}

float crossfadeGain(const Crossfade& fade, float value, bool controllerScale)
{
    const auto top = [controllerScale](int high) { return controllerScale && high == 127 ? 126 : high; };
    const int inHigh = top(fade.inHigh), outHigh = top(fade.outHigh);
    float up = 1, down = 1;
    if (value < float(fade.inLow))
        up = 0;
    else if (inHigh > fade.inLow)
        up = std::min(1.f, (value - float(fade.inLow)) / float(inHigh - fade.inLow));
    if (outHigh > fade.outLow)
        down = std::clamp((float(outHigh) - value) / float(outHigh - fade.outLow), 0.f, 1.f);
    return fade.power ? std::sqrt(up * down) : up * down;
}

// Pitch bend ramps within sfizz_render's 1024-frame blocks (plan.cpp:114-134).
std::vector<Bend> bendTimeline(const Midi& midi)
{
    std::vector<Bend> bends;
    int lastFrame = 0, lastValue = 0;
    for (const auto& e : midi.events) {
        const bool reset = e.kind == EventKind::control && e.a == 121;
        if (e.kind != EventKind::bend && !reset)
            continue;
        const int value = reset ? 0 : e.a;
        const int begin = std::max(lastFrame, e.frame / kBendBlock * kBendBlock);
        if (lastValue != value) {
            bends.push_back({begin, e.frame, lastValue, value});
            bends.push_back({e.frame, e.frame, value, value});
        }
        lastFrame = e.frame;
        lastValue = value;
    }
    return bends;
}

// Controller automation track: block-aligned linear ramps to each event, then
// optional one-pole-like smoothing evaluated in 16-frame groups (plan.cpp:136-216).
template<class Transform>
Track automation(const Midi& midi, int cc, float initial, int smoothFrames, Transform transform, bool settle)
{
    float previous = transform(initial);
    Track points{{0, previous, 0}};
    struct Ramp {
        int first, last;
        float from, to;
    };
    std::vector<Ramp> ramps;
    int lastFrame = 0;
    for (const auto& e : midi.events) {
        float value;
        if (e.kind == EventKind::control && e.a == 121) value = initial;
        else if (cc == 128 && e.kind == EventKind::bend) value = float(e.a);
        else if (e.kind == EventKind::control && e.a == cc) value = float(e.b);
        else continue;
        const float target = transform(value);
        if (settle && e.frame == 0) {
            points.front().value = target;
        } else if (target != previous) {
            ramps.push_back({std::max(lastFrame, e.frame / kBendBlock * kBendBlock), e.frame, previous, target});
        }
        previous = target;
        lastFrame = e.frame;
    }
    const auto add = [&](int frame, float value, float step) {
        auto& back = points.back();
        if (back.frame == frame)
            back = {frame, value, step};
        else if (back.step != step || std::abs(back.value + float(frame - back.frame) * step - value) > 1e-6f)
            points.push_back({frame, value, step});
    };
    if (smoothFrames < 2) {
        for (const auto& ramp : ramps) {
            if (ramp.first < ramp.last)
                add(ramp.first, ramp.from, (ramp.to - ramp.from) / float(ramp.last - ramp.first));
            add(ramp.last, ramp.to, 0);
        }
        return points;
    }
    constexpr int group = 16;
    const int smooth = std::max(group, smoothFrames);
    float current = points.front().value, target = current, step = 0;
    size_t next = 0;
    for (int frame = 0; next < ramps.size() || current != target; frame += group) {
        if (current == target && next < ramps.size() && frame < ramps[next].first)
            frame = std::max(frame, ramps[next].first / group * group);
        float goal = target;
        const int end = frame + group - 1;
        while (next < ramps.size() && end >= ramps[next].last)
            goal = ramps[next++].to;
        if (next < ramps.size() && end >= ramps[next].first) {
            const auto& ramp = ramps[next];
            goal = ramp.from + (ramp.to - ramp.from) * (float(end - ramp.first) / float(ramp.last - ramp.first));
        }
        if (goal != target) {
            target = goal;
            step = (target - current) / float(smooth);
        }
        if (current == target || step == 0) {
            add(frame, current, 0);
            continue;
        }
        add(frame, current + step, step);
        const float remaining = std::abs((target - current) / step);
        if (remaining <= group) {
            add(frame + std::max(0, int(std::ceil(remaining)) - 1), target, 0);
            current = target;
        } else {
            current += float(group) * step;
        }
    }
    return points;
}

// Host twin of the device envelope (plan.cpp:218-235), float arithmetic.
template<class E>
float envelopeLevel(const E& v, int frame)
{
    if (v.release_start >= 0 && frame >= v.release_start) {
        if (frame >= v.fade_start)
            return std::max(0.f, v.fade_value - float(frame - v.fade_start + 1) * v.fade_step);
        return v.release_value * std::exp(float(frame - v.release_start + 1) * v.release_rate);
    }
    if (frame < v.attack_start)
        return v.start;
    if (frame < v.hold_start)
        return std::min(1.f, v.start + float(frame - v.attack_start + 1) * v.attack_step);
    if (frame < v.decay_start)
        return 1;
    if (v.decay_rate == 0)
        return v.sustain;
    return std::max(v.sustain, std::exp(float(frame - v.decay_start + 1) * v.decay_rate));
}

// Exponential release to 1e-4, then a linear 50 ms fade (plan.cpp:237-248).
template<class E>
void beginRelease(E& e, int frame, int rate, float level, float seconds)
{
    e.release_start = frame;
    e.release_value = level;
    e.release_rate = -9 / (std::clamp(seconds, .006f, 100.f) * float(rate));
    const int duration = level > kSilence ? int(std::log(kSilence / level) / e.release_rate) : 0;
    e.fade_start = frame + duration;
    e.fade_value = level * std::exp(e.release_rate * float(duration));
    e.fade_step = std::max(kSilence, e.fade_value) / (float(rate) * .05f);
}

double travel(const clay_voice_span& v, int frames)
{
    if (v.pitch_slope == 0)
        return v.step * v.minimum_pitch_ratio * frames;
    return v.step * v.minimum_pitch_ratio * std::expm1(v.pitch_slope * frames) / std::expm1(v.pitch_slope);
}

double wrapLoop(const clay_voice_span& v, double position)
{
    if (v.loop_end > 0 && position >= v.loop_end)
        return v.loop_first + std::fmod(position - v.loop_first, double(v.loop_end - v.loop_first));
    return position;
}

// Frame where an unlooped sample runs out, bounded by the envelope end (plan.cpp:264-280).
int voiceEnd(const clay_voice_span& v, int envelopeEnd)
{
    if (v.loop_end > 0 || v.generator != CLAY_GEN_NONE)
        return envelopeEnd;
    const int rampFrames = std::max(0, v.bend_end - v.begin);
    const double rampTravel = travel(v, rampFrames);
    const double remaining = v.source_end - v.position;
    const double step = v.step * v.minimum_pitch_ratio;
    const double length = rampFrames && v.pitch_slope != 0 && remaining <= rampTravel
        ? std::ceil(std::log1p(std::max(0., remaining) * std::expm1(v.pitch_slope) / step) / v.pitch_slope)
        : rampFrames + std::ceil((remaining - rampTravel) / (step * std::exp(v.pitch_slope * rampFrames)));
    if (length > double(kForever - v.begin))
        return envelopeEnd;
    return std::min(envelopeEnd, v.begin + std::max(0, int(length)));
}

class Planner {
public:
    Planner(const Instrument& instrument, AudioBank& bank, int rate, unsigned seed, int limit, TailPolicy tail)
        : inst_(instrument), bank_(bank), rate_(rate), limit_(limit), tail_(tail), cc_(instrument.initialCC),
          sequence_(instrument.regions.size(), 0u), random_(seed), variation_(seed ^ 0x85ebca6bu),
          currentSwitch_(instrument.initialSwitch)
    {
    }

    Plan render(const Midi& midi)
    {
        midi_ = &midi;
        hasBend_ = std::any_of(midi.events.begin(), midi.events.end(),
                               [](const Event& e) { return e.kind == EventKind::bend && e.a != 0; });
        for (int cc : {7, 11, 10}) {
            if (inst_.usedCC[size_t(cc)]) {
                plan_.controls.push_back({{0, cc == 10 ? 0.f : 1.f, 0.f}});
            } else {
                const int curve = cc == 10 ? 1 : 4;
                plan_.controls.push_back(automation(midi, cc, inst_.initialCC[size_t(cc)], rate_ / 100,
                    [&](float v) { return curveValue(inst_, curve, v); }, natural()));
            }
        }
        const auto bends = bendTimeline(midi);
        size_t nextBend = 0;
        for (const auto& e : midi.events) {
            while (nextBend < bends.size() && bends[nextBend].begin <= e.frame)
                changeBend(bends[nextBend++]);
            prune(e.frame);
            switch (e.kind) {
            case EventKind::on: noteOn(e); break;
            case EventKind::off: noteOff(e); break;
            case EventKind::control: control(e); break;
            case EventKind::bend: break;
            }
        }
        prune(midi.end);
        for (auto& voice : active_) {
            if (!natural() && voice.kind == VoiceKind::attack && voice.region->loop != Loop::oneShot)
                release(voice, midi.end);
            if (voice.span.end == kForever && limit_ == kForever)
                throw std::runtime_error("a looping one-shot or release sample has no finite end");
            emit(voice, voice.span.end);
        }
        if (!natural())
            plan_.frames = std::max(plan_.frames, midi.end);
        if (plan_.spans.empty())
            throw std::runtime_error("no playable notes for this SFZ and MIDI");
        buildLfoTracks();
        return std::move(plan_);
    }

private:
    bool natural() const { return tail_ == TailPolicy::natural; }

    void emit(const Voice& voice, int end)
    {
        auto span = voice.span;
        span.end = std::min(span.end, end);
        plan_.tailClipped |= span.end > limit_ && span.begin < span.end;
        span.end = std::min(span.end, limit_);
        if (span.begin < span.end) {
            for (int i = voice.firstLfo; i < voice.firstLfo + voice.lfoCount; ++i)
                lfos_[size_t(i)].end = std::max(lfos_[size_t(i)].end, span.end);
            plan_.frames = std::max(plan_.frames, span.end);
            plan_.spans.push_back(span);
        }
    }

    void prune(int frame)
    {
        auto keep = std::remove_if(active_.begin(), active_.end(), [&](const Voice& v) {
            if (v.span.end > frame)
                return false;
            emit(v, v.span.end);
            return true;
        });
        active_.erase(keep, active_.end());
    }

    void split(Voice& voice, int frame, double position)
    {
        emit(voice, frame);
        voice.span.begin = frame;
        voice.span.position = position;
    }

    bool pedal(const Region& r) const { return r.sustainEnabled && cc_[size_t(r.sustainCC)] >= float(r.sustainLow); }

    bool switchesMatch(const Region& r) const
    {
        return (r.switchLow < 0 || (currentSwitch_ >= r.switchLow && currentSwitch_ <= r.switchHigh)) &&
               (r.switchDown < 0 || down_[size_t(r.switchDown)]) && (r.switchUp < 0 || !down_[size_t(r.switchUp)]) &&
               (r.switchPrevious < 0 || previousNote_ == r.switchPrevious);
    }

    bool controllersMatch(const Region& r) const
    {
        return std::all_of(r.conditions.begin(), r.conditions.end(), [&](const Condition& c) {
            return cc_[c.cc] >= float(c.low) && cc_[c.cc] <= float(c.high);
        });
    }

    bool matches(const Region& r, int velocity, float draw) const
    {
        return velocity >= r.lowVelocity && velocity <= r.highVelocity && draw >= r.randomLow && draw < r.randomHigh &&
               switchesMatch(r) && controllersMatch(r);
    }

    // Plain pitch bend: a ramp span with exponential step change, or a constant step (plan.cpp:396-420).
    void applyBend(Voice& voice) const
    {
        auto& s = voice.span;
        if (voice.region->bendSmooth > 0) {
            s.bend_end = s.begin;
            s.pitch_slope = 0;
            s.step = voice.baseStep;
            return;
        }
        const auto cents = [&](int value) {
            const double normalized = std::max(-8191, value) / 8191.0;
            return normalized >= 0 ? normalized * voice.region->bendUp : -normalized * voice.region->bendDown;
        };
        const double to = cents(bend_.to);
        s.bend_end = bend_.end;
        s.pitch_slope = 0;
        if (s.begin < bend_.end) {
            const double from = cents(bend_.from);
            const double slope = (to - from) / (bend_.end - bend_.begin);
            s.pitch_slope = slope * (std::log(2.) / 1200);
            s.step = voice.baseStep * std::exp2((from + (s.begin - bend_.begin) * slope) / 1200);
        } else {
            s.step = voice.baseStep * std::exp2(to / 1200);
        }
    }

    void changeBend(const Bend& bend)
    {
        prune(bend.begin);
        bend_ = bend;
        for (auto& voice : active_) {
            if (voice.region->bendSmooth > 0)
                continue;
            const int start = std::max(bend.begin, voice.span.begin);
            const double phase = voice.span.position + travel(voice.span, start - voice.span.begin);
            split(voice, start, phase);
            applyBend(voice);
            voice.span.end = voiceEnd(voice.span, fadeEnd(voice.span));
        }
    }

    int fadeEnd(const clay_voice_span& s) const { return s.release_start < 0 ? kForever : s.fade_start + rate_ / 20 + 1; }

    float controllerValue(const Modulation& m, int key, int velocity, VoiceKind kind) const
    {
        if (m.cc == 140)
            return keyDelta_;
        if (m.cc == 141)
            return std::abs(keyDelta_);
        float value = 0;
        if (m.cc < 128) value = cc_[size_t(m.cc)];
        else if (m.cc == 133) value = float(key);
        else if ((m.cc == 131 && kind == VoiceKind::attack) || (m.cc == 132 && kind == VoiceKind::release)) value = float(velocity);
        else if (m.cc == 134) value = keysDown_ > 0 ? 127.f : 0.f;
        else if (m.cc == 135) value = 127 * randomValue_;
        else if (m.cc == 136) value = 127 * (2 * randomValue_ - 1);
        else if (m.cc == 137) value = alternate_ ? 127.f : 0.f;
        else if (const auto it = inst_.extendedCCDefaults.find(m.cc); it != inst_.extendedCCDefaults.end()) value = it->second;
        return curveValue(inst_, m.curve, value);
    }

    int controllerTrack(const Modulation& m)
    {
        const auto key = std::make_tuple(m.cc, m.curve, m.smooth);
        const auto [it, inserted] = tracks_.emplace(key, int(plan_.controls.size()));
        if (inserted)
            plan_.controls.push_back(automation(*midi_, m.cc, inst_.initialCC[size_t(m.cc)], m.smooth * rate_ / 1000,
                [&](float v) { return curveValue(inst_, m.curve, v); }, natural()));
        return it->second;
    }

    int smoothBendTrack(const Region& r)
    {
        const auto key = std::make_tuple(r.bendUp, r.bendDown, r.bendSmooth);
        const auto [it, inserted] = bendTracks_.emplace(key, int(plan_.controls.size()));
        if (inserted)
            plan_.controls.push_back(automation(*midi_, 128, 0, r.bendSmooth * rate_ / 1000, [&](float v) {
                const float bend = std::max(-8191.f, v) / 8191;
                return bend >= 0 ? bend * r.bendUp : -bend * r.bendDown;
            }, natural()));
        return it->second;
    }

    int controllerFadeTrack(int cc, const Crossfade& fade)
    {
        const auto key = std::make_tuple(cc, fade.inLow, fade.inHigh, fade.outLow, fade.outHigh, fade.power);
        const auto [it, inserted] = fadeTracks_.emplace(key, int(plan_.controls.size()));
        if (inserted)
            plan_.controls.push_back(automation(*midi_, cc, inst_.initialCC[size_t(cc)], rate_ / 200,
                [&](float v) { return crossfadeGain(fade, v, true); }, natural()));
        return it->second;
    }

    float param(const Region& r, ModTarget target, float base, int key, int velocity, VoiceKind kind) const
    {
        if (target >= ModTarget::envDelay && target <= ModTarget::envRelease)
            base += r.velocityEnvelope[size_t(int(target) - int(ModTarget::envDelay))] * (float(velocity) / 127.f);
        for (const auto& m : r.modulation)
            if (m.target == target)
                base += m.depth * controllerValue(m, key, velocity, kind);
        return base;
    }

    // Filter EG (slot 0) and pitch EG (slot 1) shapes (plan.cpp:507-536).
    void initSecondary(Voice& voice)
    {
        auto& s = voice.span;
        for (int slot = 0; slot < 2; ++slot) {
            const auto& src = voice.region->secondary[size_t(slot)];
            if (!src.enabled)
                continue;
            const auto first = slot == 0 ? ModTarget::filterEnvDelay : ModTarget::pitchEnvDelay;
            std::array<float, 7> values;
            for (size_t p = 0; p < values.size(); ++p) {
                const float base = src.values[p] + src.velocity[p] * (float(voice.velocity) / 127.f);
                const float value = param(*voice.region, static_cast<ModTarget>(int(first) + int(p)), base, voice.key,
                                          voice.velocity, voice.kind);
                values[p] = std::clamp(value, 0.f, p == 4 || p == 5 ? 1.f : 100.f);
            }
            auto& e = s.extra_envelope[slot];
            s.active_envelope[slot] = 1;
            e.start = values[5];
            e.attack_start = s.begin + int(values[0] * float(rate_));
            const float attack = values[1] * float(rate_);
            e.attack_step = attack > 0 ? 1 / attack : 1;
            e.hold_start = e.attack_start + std::max(0, int(std::ceil((1 - e.start) * attack)) - 1);
            e.decay_start = e.hold_start + int(values[2] * float(rate_));
            e.decay_rate = values[3] > 0 ? -9 / (std::max(.006f, values[3]) * float(rate_)) : 0;
            e.sustain = values[4];
            e.release_start = -1;
            e.release_seconds = values[6];
            e.depth = src.depth + src.velocityDepth * (float(voice.velocity) / 127.f);
        }
    }

    float appendLfos(Voice& voice, int frame, bool& modulatedStereo)
    {
        voice.firstLfo = int(lfos_.size());
        float pitchBound = 0;
        for (const auto& entry : voice.region->lfos) {
            const auto& lfo = entry.second;
            LfoTrack track{&lfo, frame, frame, 0, 0, {}, {}};
            float delay = lfo.delay, fade = lfo.fade;
            auto bounds = lfo.depths;
            for (auto& d : bounds)
                d.second = std::abs(d.second);
            for (const auto& m : lfo.modulation) {
                const float value = controllerValue(m, voice.key, voice.velocity, voice.kind);
                if (m.target == ModTarget::lfoDelay) {
                    delay += m.depth * value;
                } else if (m.target == ModTarget::lfoFade) {
                    fade += m.depth * value;
                } else {
                    const int source = m.cc < 128 ? controllerTrack(m) : -1;
                    track.inputs.push_back({source, m.target, source < 0 ? m.depth * value : m.depth});
                    if (m.target < ModTarget::lfoFrequency || m.target >= ModTarget::eq1Frequency) {
                        float maximum = std::abs(value);
                        if (source >= 0)
                            for (const auto& p : plan_.controls[size_t(source)])
                                maximum = std::max(maximum, std::abs(p.value));
                        bounds[m.target] += std::abs(m.depth) * maximum;
                    }
                }
            }
            track.delay = int(std::clamp(delay, 0.f, 100.f) * float(rate_));
            track.fade = int(std::clamp(fade, 0.f, 100.f) * float(rate_));
            for (const auto& [target, bound] : bounds) {
                if (bound == 0)
                    continue;
                modulatedStereo |= target == ModTarget::width || target == ModTarget::position;
                if (target == ModTarget::pitch)
                    pitchBound += bound;
                const int index = int(plan_.controls.size());
                plan_.controls.emplace_back();
                track.outputs.emplace_back(target, index);
                plan_.modulation.push_back({index, int(target), 1.f, -1});
            }
            if (!track.outputs.empty())
                lfos_.push_back(std::move(track));
        }
        voice.lfoCount = int(lfos_.size()) - voice.firstLfo;
        return pitchBound;
    }

    // LFO output tracks, sampled every 16 frames for triangle/sine (plan.cpp:589-639).
    void buildLfoTracks()
    {
        std::mt19937 sampleHold(1);
        for (const auto& lfo : lfos_) {
            const auto& src = *lfo.source;
            const int start = lfo.begin + lfo.delay;
            const int stride = src.wave <= 1 ? 16 : 1;
            double phase = 0;
            int cycle = -1;
            float held = 0;
            std::array<float, int(ModTarget::eq3Gain) + 1> baseDepth{};
            for (const auto& d : src.depths)
                baseDepth[size_t(d.first)] = d.second;
            for (const auto& out : lfo.outputs)
                plan_.controls[size_t(out.second)].push_back({0, 0, 0});
            for (int frame = start; frame < lfo.end; frame += stride) {
                float frequency = src.frequency, shift = src.phase;
                auto depth = baseDepth;
                for (const auto& in : lfo.inputs) {
                    const float v = in.track < 0 ? in.depth : in.depth * trackAt(plan_.controls[size_t(in.track)], frame);
                    if (in.target == ModTarget::lfoFrequency) frequency += v;
                    else if (in.target == ModTarget::lfoPhase) shift += v;
                    else depth[size_t(in.target)] += v;
                }
                float wave = lfoWave(src.wave, phase + shift);
                if (src.wave == 12) {
                    const int nextCycle = int(phase);
                    if (cycle != nextCycle) {
                        held = 2 * std::generate_canonical<float, 24>(sampleHold) - 1;
                        cycle = nextCycle;
                    }
                    wave = held;
                }
                const float fade = lfo.fade == 0 ? 1.f : std::min(1.f, float(frame - start) / float(lfo.fade));
                for (const auto& out : lfo.outputs) {
                    auto& points = plan_.controls[size_t(out.second)];
                    const float value = depth[size_t(out.first)] * wave * fade;
                    if (points.back().frame == frame) {
                        points.back().value = value;
                    } else {
                        if (stride > 1 && frame - points.back().frame == stride)
                            points.back().step = (value - points.back().value) / float(stride);
                        if (stride > 1 || value != points.back().value)
                            points.push_back({frame, value, 0});
                    }
                }
                phase += double(std::clamp(frequency, 0.f, 1000.f)) * stride / rate_;
            }
        }
    }

    // egN pitch envelopes as control tracks plus pitch bindings (plan.cpp:641-691).
    float appendPitchEnvelopes(Voice& voice)
    {
        float bound = 0;
        for (const auto& entry : voice.region->pitchEnvelopes) {
            const auto& src = entry.second;
            std::vector<clay_modulation> bindings;
            if (src.depth != 0) {
                bindings.push_back({-1, int(ModTarget::pitch), src.depth, -1});
                bound += std::abs(src.depth);
            }
            for (const auto& [cc, depth] : src.depthCC) {
                const Modulation m{ModTarget::pitch, cc, depth};
                const float value = controllerValue(m, voice.key, voice.velocity, voice.kind);
                if (cc < 128) {
                    bindings.push_back({controllerTrack(m), int(ModTarget::pitch), depth, -1});
                    bound += std::abs(depth);
                } else if (value != 0) {
                    bindings.push_back({-1, int(ModTarget::pitch), depth * value, -1});
                    bound += std::abs(depth * value);
                }
            }
            if (bindings.empty())
                continue;
            const int shape = int(plan_.controls.size());
            voice.flexShapes.push_back(shape);
            plan_.controls.emplace_back();
            auto& track = plan_.controls.back();
            track.push_back({0, 0, 0});
            const auto write = [&](int frame, float level, float step) {
                if (track.back().frame == frame) track.back() = {frame, level, step};
                else track.push_back({frame, level, step});
            };
            int frame = voice.triggerFrame;
            float level = 0;
            for (const auto& [index, point] : src.points) {
                float time = point.seconds;
                for (const auto& [cc, amount] : point.timeCC)
                    time += amount * controllerValue({ModTarget::pitch, cc}, voice.key, voice.velocity, voice.kind);
                const int frames = int(std::clamp(time, 0.f, 100.f) * float(rate_));
                if (frames > 0)
                    write(frame, level, (point.level - level) / float(frames));
                level = point.level;
                frame += frames;
                write(frame, level, 0);
            }
            for (auto b : bindings) {
                b.shape_track = shape;
                plan_.modulation.push_back(b);
            }
        }
        return bound;
    }

    // Start a voice (plan.cpp:693-830).
    void trigger(const Region& r, int frame, int key, int velocity, VoiceKind kind, int duration = 0, int controller = -1)
    {
        if (inst_.hasOffBy)
            choke(r, frame, controller >= 0 ? controller : key, kind);
        if (r.notePolyphony)
            limitNoteVoices(r, frame, key, velocity);
        const Sample* sample = r.generator == Generator::none ? &bank_.decode(r.sample) : nullptr;
        const auto vary = [&](float extent) {
            return extent == 0 ? 0.f : extent * std::generate_canonical<float, 24>(variation_);
        };
        Voice voice{};
        voice.key = key;
        voice.velocity = velocity;
        voice.triggerFrame = frame;
        voice.region = &r;
        voice.kind = kind;
        const auto modulated = [&](ModTarget t, float base) { return param(r, t, base, key, velocity, kind); };
        const auto seconds = [&](ModTarget t, float base) { return std::clamp(modulated(t, base), 0.f, 100.f); };
        voice.releaseSeconds = seconds(ModTarget::envRelease, r.release);
        const double baseRate = sample ? sample->rate : 440 * std::exp2((r.rootKey - 69) / 12.);
        voice.baseStep = (baseRate / rate_) *
            std::exp2((r.keyTrack * float(key - r.rootKey) + r.tune + float(100 * r.transpose) + vary(r.pitchRandom)) / 1200.0);
        auto& s = voice.span;
        s.sample = sample ? sample->offset : 0;
        s.channels = sample ? sample->channels : 1;
        s.generator = int(r.generator);
        if (r.generator == Generator::noise) {
            s.noise_seed = variation_();
            s.noise_start = frame;
        }
        s.begin = frame + int(seconds(ModTarget::delay, r.delay) * float(rate_));
        s.source_end = r.end;
        const int randomOffset = r.offsetRandom ? std::uniform_int_distribution<int>(0, r.offsetRandom)(variation_) : 0;
        if (sample) {
            s.position = std::trunc(std::clamp(double(modulated(ModTarget::offset, float(r.offset))) + randomOffset, 0.,
                                               double(r.end)));
        } else {
            const double phase = r.oscillatorPhase >= 0 ? r.oscillatorPhase : randomValue_;
            s.position = phase - std::floor(phase);
        }
        applyBend(voice);
        if (sample && (r.loop == Loop::continuous || r.loop == Loop::sustain)) {
            s.loop_first = r.loopStart;
            s.loop_end = r.loopEnd;
            s.loop_crossfade = int(std::lround(std::min<double>(r.loopStart, double(r.loopCrossfade) * sample->rate)));
        }
        s.start = std::clamp(modulated(ModTarget::envStart, r.start), 0.f, 1.f);
        s.attack_start = s.begin + int(seconds(ModTarget::envDelay, r.envelopeDelay) * float(rate_));
        const float attackFrames = seconds(ModTarget::envAttack, r.attack) * float(rate_);
        s.attack_step = attackFrames > 0 ? 1 / attackFrames : 1;
        s.hold_start = s.attack_start + std::max(0, int(std::ceil((1 - s.start) * attackFrames)) - 1);
        s.decay_start = s.hold_start + int(seconds(ModTarget::envHold, r.hold) * float(rate_));
        const float decay = seconds(ModTarget::envDecay, r.decay);
        s.decay_rate = decay > 0 ? -9 / (std::max(.006f, decay) * float(rate_)) : 0;
        s.sustain = std::clamp(modulated(ModTarget::envSustain, r.sustain), 0.f, 1.f);
        s.release_start = -1;
        initSecondary(voice);
        s.filter_track = -1;
        s.sample_rate = rate_;
        for (size_t i = 0; i < r.filters.size(); ++i) {
            const auto& f = r.filters[i];
            if (f.type == FilterType::none)
                continue;
            if (s.filter_track < 0)
                s.filter_track = plan_.filterTracks++;
            const auto velocityTarget = i == 0 ? ModTarget::filterVelocity : ModTarget::filterVelocity2;
            const float cents = f.keyTrack * float(key - f.keyCenter) + vary(f.random) +
                                modulated(velocityTarget, f.velocityTrack) * float(velocity) / 127.f;
            s.filters[i] = {int(f.type), std::clamp(f.cutoff * std::exp2(std::clamp(cents, -24000.f, 24000.f) / 1200), 0.f, 20000.f),
                            f.resonance, 0, 0};
        }
        static const float eqDefaults[3] = {50, 500, 5000};
        for (size_t i = 0; i < r.eq.size(); ++i) {
            const auto& eq = r.eq[i];
            if (!eq.enabled)
                continue;
            if (s.filter_track < 0)
                s.filter_track = plan_.filterTracks++;
            s.filters[i + 2] = {int(FilterType::peak), eq.frequency == 0 ? eqDefaults[i] : eq.frequency, 0, eq.bandwidth, eq.gain};
        }
        s.first_modulation = int(plan_.modulation.size());
        bool modulatedStereo = false;
        float pitchBound = std::abs(s.extra_envelope[1].depth);
        for (const auto& m : r.modulation) {
            if ((m.target >= ModTarget::envDelay && m.target < ModTarget::cutoff) ||
                (m.target >= ModTarget::filterVelocity && m.target < ModTarget::filterEnvDepth))
                continue;
            modulatedStereo |= m.target == ModTarget::width || m.target == ModTarget::position;
            if (m.target == ModTarget::pitch || m.target == ModTarget::pitchEnvDepth)
                pitchBound += std::abs(m.cc < 128 ? m.depth : m.depth * controllerValue(m, key, velocity, kind));
            if (m.cc < 128)
                plan_.modulation.push_back({controllerTrack(m), int(m.target), m.depth, -1});
            else
                plan_.modulation.push_back({-1, int(m.target), m.depth * controllerValue(m, key, velocity, kind), -1});
        }
        for (const auto& [cc, fade] : r.controllerFades)
            plan_.modulation.push_back({controllerFadeTrack(cc, fade), int(ModTarget::amplitude), 1.f, -1});
        if (r.bendSmooth > 0 && hasBend_) {
            pitchBound += std::max(std::abs(r.bendUp), std::abs(r.bendDown));
            plan_.modulation.push_back({smoothBendTrack(r), int(ModTarget::pitch), 1.f, -1});
        }
        pitchBound += appendLfos(voice, frame, modulatedStereo);
        pitchBound += appendPitchEnvelopes(voice);
        s.minimum_pitch_ratio = std::exp2(-std::min(24000.f, pitchBound) / 1200);
        if (pitchBound > 0) {
            s.variable_pitch = 1;
            if (s.filter_track < 0)
                s.filter_track = plan_.filterTracks++;
        }
        s.modulation_count = int(plan_.modulation.size()) - s.first_modulation;
        const float noteGain = kind == VoiceKind::controller ? 1.f :
            velocityGain(r, velocity) * crossfadeGain(r.keyFade, float(key), false) *
            crossfadeGain(r.velocityFade, float(velocity), true);
        const float gain = r.amplitude * noteGain *
            std::pow(10.f, ((natural() ? 0 : kMasterGainDb) + r.volume + vary(r.ampRandom) -
                            r.releaseDecay * float(duration) / float(rate_) + 3) / 20);
        constexpr float panCompensation = 1.4125375446227544f;
        const bool stereoMatrix = s.channels == 2 && (modulatedStereo || r.width != 1 || r.position != 0);
        s.stereo_matrix = stereoMatrix;
        const float positionGain = s.channels == 1 ? 1.f : .99881488f;
        s.gain = gain * (stereoMatrix ? panCompensation : positionGain);
        s.pan = r.pan;
        s.width = r.width;
        s.stereo_position = r.position;
        s.left = s.gain * std::cos((1 + r.pan) * kQuarterPi);
        s.right = s.gain * std::cos((1 - r.pan) * kQuarterPi);
        s.width_direct = stereoMatrix ? std::cos((1 - r.width) * kQuarterPi) : 1.f;
        s.width_cross = stereoMatrix ? std::cos((1 + r.width) * kQuarterPi) : 0.f;
        s.position_left = stereoMatrix ? std::cos((1 + r.position) * kQuarterPi) : 1.f;
        s.position_right = stereoMatrix ? std::cos((1 - r.position) * kQuarterPi) : 1.f;
        s.end = voiceEnd(s, kForever);
        active_.push_back(std::move(voice));
    }

    void release(Voice& voice, int frame)
    {
        if (voice.span.release_start >= 0 || voice.kind != VoiceKind::attack || voice.region->loop == Loop::oneShot)
            return;
        if (voice.region->envelopeDynamic)
            voice.releaseSeconds = param(*voice.region, ModTarget::envRelease, voice.region->release, voice.key,
                                         voice.velocity, voice.kind);
        startRelease(voice, frame, voice.releaseSeconds);
    }

    // plan.cpp:842-873
    void startRelease(Voice& voice, int frame, float seconds)
    {
        auto& s = voice.span;
        if (frame >= s.end)
            return;
        if (frame < s.begin) {
            s.end = s.begin;
            return;
        }
        const float level = envelopeLevel(s, frame - 1);
        const bool exitLoop = voice.region->loop == Loop::sustain && s.loop_end > 0;
        if (exitLoop || s.release_start >= 0) {
            const double phase = s.position + travel(s, frame - s.begin);
            split(voice, frame, exitLoop ? wrapLoop(s, phase) : phase);
            applyBend(voice);
            if (exitLoop)
                s.loop_end = 0;
        }
        beginRelease(s, frame, rate_, level, seconds);
        for (int shape : voice.flexShapes) {
            auto& track = plan_.controls[size_t(shape)];
            track.erase(std::lower_bound(track.begin(), track.end(), frame,
                                         [](const clay_control_point& p, int f) { return p.frame < f; }),
                        track.end());
            track.push_back({frame, 0, 0});
        }
        for (int slot = 0; slot < 2; ++slot) {
            auto& e = s.extra_envelope[slot];
            if (s.active_envelope[slot] && e.release_start < 0)
                beginRelease(e, frame, rate_, envelopeLevel(e, frame - 1), e.release_seconds);
        }
        s.end = voiceEnd(s, fadeEnd(s));
    }

    void choke(const Region& incoming, int frame, int key, VoiceKind cause)
    {
        std::vector<std::pair<int, int>> released;
        for (auto& voice : active_) {
            const auto& r = *voice.region;
            if (voice.span.end <= frame || voice.kind == VoiceKind::release || voice.choked || !r.chokedByGroup ||
                r.offBy != incoming.group || (r.group == incoming.group && voice.key == key && controllersMatch(r)))
                continue;
            forceOff(voice, frame);
            if (cause != VoiceKind::controller && voice.kind == VoiceKind::attack)
                released.emplace_back(voice.key, voice.velocity);
        }
        for (const auto& [k, v] : released)
            releaseRegions(frame, k, v, std::max(0, frame - noteStart_[size_t(k)]));
    }

    void forceOff(Voice& voice, int frame)
    {
        const auto& r = *voice.region;
        voice.choked = true;
        const float time = r.offMode == OffMode::normal ? voice.releaseSeconds : r.offMode == OffMode::fast ? .006f : r.offTime;
        startRelease(voice, frame, time);
    }

    void limitNoteVoices(const Region& region, int frame, int key, int velocity)
    {
        std::vector<Voice*> candidates;
        int count = 0;
        for (auto& voice : active_) {
            if (voice.choked || voice.span.end <= frame || voice.key != key || voice.region->group != region.group)
                continue;
            ++count;
            if (!region.selfMask || voice.velocity <= velocity)
                candidates.push_back(&voice);
        }
        if (count < region.notePolyphony)
            return;
        std::stable_sort(candidates.begin(), candidates.end(), [&](const Voice* a, const Voice* b) {
            return region.selfMask ? a->velocity < b->velocity : a->triggerFrame < b->triggerFrame;
        });
        for (auto* candidate : candidates) {
            if (candidate->choked)
                continue;
            for (auto& sister : active_)
                if (!sister.choked && sister.key == candidate->key && sister.kind == candidate->kind &&
                    sister.triggerFrame == candidate->triggerFrame)
                    forceOff(sister, frame);
            if (--count < region.notePolyphony)
                break;
        }
    }

    void noteOn(const Event& e)
    {
        const bool switching = inst_.switchKeys[size_t(e.a)];
        if (switching)
            currentSwitch_ = e.a;
        if (!down_[size_t(e.a)] && !switching)
            ++keysDown_;
        if (!switching) {
            keyDelta_ = lastNote_ < 0 ? 0.f : float(e.a - lastNote_);
            lastNote_ = e.a;
            randomValue_ = std::generate_canonical<float, 24>(extendedRandom_);
            alternate_ = !alternate_;
        }
        down_[size_t(e.a)] = true;
        noteVelocity_[size_t(e.a)] = e.b;
        noteStart_[size_t(e.a)] = e.frame;
        const float draw = std::generate_canonical<float, 24>(random_);
        for (const int index : inst_.byKey[size_t(e.a)]) {
            const auto& r = inst_.regions[size_t(index)];
            const bool step = sequence_[size_t(index)]++ % unsigned(r.sequenceLength) == unsigned(r.sequencePosition - 1);
            if (!step || !matches(r, e.b, draw))
                continue;
            if (r.trigger == Trigger::attack || (r.trigger == Trigger::first && keysDown_ == 1) ||
                (r.trigger == Trigger::legato && keysDown_ > 1))
                trigger(r, e.frame, e.a, e.b, VoiceKind::attack);
        }
        previousNote_ = e.a;
    }

    bool hasAttackVoice(const Region& r) const
    {
        return std::any_of(active_.begin(), active_.end(), [&](const Voice& v) {
            return v.kind == VoiceKind::attack && v.key >= r.lowKey && v.key <= r.highKey && v.velocity >= r.lowVelocity &&
                   v.velocity <= r.highVelocity;
        });
    }

    void noteOff(const Event& e)
    {
        if (!down_[size_t(e.a)])
            return;
        down_[size_t(e.a)] = false;
        if (!inst_.switchKeys[size_t(e.a)])
            --keysDown_;
        for (auto& voice : active_) {
            if (voice.key == e.a && voice.kind == VoiceKind::attack) {
                voice.noteOff = true;
                if (!pedal(*voice.region) && !voice.sostenuto)
                    release(voice, e.frame);
            }
        }
        releaseRegions(e.frame, e.a, noteVelocity_[size_t(e.a)], e.frame - noteStart_[size_t(e.a)]);
    }

    void releaseRegions(int frame, int key, int velocity, int duration)
    {
        const float draw = std::generate_canonical<float, 24>(random_);
        for (const int index : inst_.byKey[size_t(key)]) {
            const auto& r = inst_.regions[size_t(index)];
            const unsigned counter = sequence_[size_t(index)];
            const bool sequenceMatches = counter == 0 ? !r.explicitSequencePosition :
                (counter - 1) % unsigned(r.sequenceLength) == unsigned(r.sequencePosition - 1);
            if ((r.trigger != Trigger::release && r.trigger != Trigger::releaseKey) || !sequenceMatches ||
                !matches(r, velocity, draw))
                continue;
            if (r.trigger == Trigger::release && (pedal(r) || sostenutoKeys_[size_t(key)]))
                delayed_.push_back({key, velocity, duration, &r, sostenutoKeys_[size_t(key)]});
            else if (r.trigger == Trigger::releaseKey || r.releaseDead || hasAttackVoice(r))
                trigger(r, frame, key, velocity, VoiceKind::release, duration);
        }
    }

    void updateDynamicEnvelopes(const std::array<float, 128>& before, int frame)
    {
        for (auto& voice : active_) {
            if (voice.choked)
                continue;
            for (const auto& m : voice.region->modulation) {
                if (m.cc >= 128 || before[size_t(m.cc)] == cc_[size_t(m.cc)] || m.depth == 0)
                    continue;
                const bool filter = m.target >= ModTarget::filterEnvDelay && m.target <= ModTarget::filterEnvRelease;
                const bool pitch = m.target >= ModTarget::pitchEnvDelay && m.target <= ModTarget::pitchEnvRelease;
                if ((filter && voice.region->secondary[0].dynamic) || (pitch && voice.region->secondary[1].dynamic))
                    throw std::runtime_error("changing a dynamic pitch or filter envelope mid-voice is unsupported");
                if (!voice.region->envelopeDynamic || m.target < ModTarget::envDelay || m.target > ModTarget::envRelease)
                    continue;
                if (m.target != ModTarget::envRelease)
                    throw std::runtime_error("changing a dynamic amplitude envelope mid-voice is unsupported for CC" +
                                             std::to_string(m.cc));
                if (voice.span.release_start >= 0 && frame < voice.span.fade_start) {
                    const float secs = param(*voice.region, ModTarget::envRelease, voice.region->release, voice.key,
                                             voice.velocity, voice.kind);
                    startRelease(voice, frame, secs);
                    break;
                }
            }
        }
    }

    void control(const Event& e)
    {
        if (e.a == 121) {
            const auto before = cc_;
            cc_ = inst_.initialCC;
            updateDynamicEnvelopes(before, e.frame);
            if (natural()) {
                setSostenuto(false);
                releasePedals(e.frame, 66, before[66]);
            }
            for (auto& voice : active_)
                if (voice.noteOff && !pedal(*voice.region))
                    release(voice, e.frame);
            return;
        }
        if (!natural() && e.a == 66 && e.b != 0)
            throw std::runtime_error("unsupported MIDI sostenuto CC66");
        if (natural() && (e.a == 120 || e.a == 123)) {
            for (const auto& voice : active_)
                emit(voice, e.frame);
            active_.clear();
            delayed_.clear();
            down_.fill(false);
            sostenutoKeys_.fill(false);
            keysDown_ = 0;
            return;
        }
        if (e.a >= 120)
            throw std::runtime_error("unsupported MIDI channel mode CC" + std::to_string(e.a));
        const bool sostenuto = natural() && e.a == 66;
        if (!inst_.usedCC[size_t(e.a)] && !sostenuto)
            return;
        const auto before = cc_;
        const float previous = before[size_t(e.a)];
        cc_[size_t(e.a)] = float(e.b);
        if (previous != cc_[size_t(e.a)])
            updateDynamicEnvelopes(before, e.frame);
        if (sostenuto)
            setSostenuto(e.b == 127);
        releasePedals(e.frame, e.a, previous);
        const auto& triggers = inst_.byCC[size_t(e.a)];
        if (triggers.empty())
            return;
        const float draw = std::generate_canonical<float, 24>(random_);
        for (const int index : triggers) {
            const auto& r = inst_.regions[size_t(index)];
            const auto c = std::find_if(r.ccTriggers.begin(), r.ccTriggers.end(), [&](const Condition& x) { return x.cc == e.a; });
            if (e.b < c->low || e.b > c->high || draw < r.randomLow || draw >= r.randomHigh)
                continue;
            const bool step = sequence_[size_t(index)]++ % unsigned(r.sequenceLength) == unsigned(r.sequencePosition - 1);
            if (step && switchesMatch(r) && previous != float(e.b) && controllersMatch(r))
                trigger(r, e.frame, r.rootKey, e.b, VoiceKind::controller, 0, e.a);
        }
    }

    void setSostenuto(bool down)
    {
        if (sostenutoDown_ == down)
            return;
        sostenutoDown_ = down;
        sostenutoKeys_ = down ? down_ : std::array<bool, 128>{};
        for (auto& voice : active_)
            voice.sostenuto = down && !voice.noteOff;
        if (!down)
            for (auto& note : delayed_)
                note.sostenuto = false;
    }

    void releasePedals(int frame, int cc, float previous)
    {
        const bool sostenutoUp = natural() && cc == 66 && !sostenutoDown_;
        for (auto& voice : active_) {
            const auto& r = *voice.region;
            const bool sustainChanged = r.sustainEnabled && r.sustainCC == cc && previous >= float(r.sustainLow);
            if (voice.noteOff && (sustainChanged || sostenutoUp) && !pedal(r) && !voice.sostenuto)
                release(voice, frame);
        }
        std::vector<DelayedRelease> due;
        auto keep = std::remove_if(delayed_.begin(), delayed_.end(), [&](const DelayedRelease& note) {
            const auto& r = *note.region;
            if ((r.sustainCC != cc && !sostenutoUp) || pedal(r) || note.sostenuto)
                return false;
            if (r.releaseDead || hasAttackVoice(r))
                due.push_back(note);
            return true;
        });
        delayed_.erase(keep, delayed_.end());
        for (const auto& note : due)
            trigger(*note.region, frame, note.key, note.velocity, VoiceKind::release, note.duration);
    }

    const Instrument& inst_;
    AudioBank& bank_;
    int rate_, limit_;
    TailPolicy tail_;
    std::array<float, 128> cc_;
    std::array<bool, 128> down_{}, sostenutoKeys_{};
    bool sostenutoDown_ = false;
    std::array<int, 128> noteStart_{}, noteVelocity_{};
    std::vector<unsigned> sequence_;
    std::mt19937 random_, variation_;
    std::mt19937 extendedRandom_{1};
    std::vector<Voice> active_;
    std::vector<DelayedRelease> delayed_;
    int keysDown_ = 0;
    int currentSwitch_;
    Bend bend_;
    const Midi* midi_ = nullptr;
    bool hasBend_ = false, alternate_ = false;
    int lastNote_ = -1, previousNote_ = -1;
    float keyDelta_ = 0, randomValue_ = 0;
    std::vector<LfoTrack> lfos_;
    std::map<std::tuple<int, int, int>, int> tracks_;
    std::map<std::tuple<float, float, int>, int> bendTracks_;
    std::map<std::tuple<int, int, int, int, int, bool>, int> fadeTracks_;
    Plan plan_;
};

} // namespace

Plan planRender(const Instrument& instrument, AudioBank& bank, const Midi& midi, int rate, unsigned seed, int limit,
                TailPolicy tail)
{
    return Planner(instrument, bank, rate, seed, limit, tail).render(midi);
}

void makeView(const Plan& plan, PlanView& out)
{
    out.points.clear();
    out.offsets.assign(1, 0);
    for (const auto& track : plan.controls) {
        out.points.insert(out.points.end(), track.begin(), track.end());
        out.offsets.push_back(int32_t(out.points.size()));
    }
    auto& v = out.view;
    v = {};
    v.spans = plan.spans.data();
    v.span_count = int32_t(plan.spans.size());
    v.points = out.points.data();
    v.track_offsets = out.offsets.data();
    v.track_count = int32_t(plan.controls.size());
    v.modulation = plan.modulation.data();
    v.modulation_count = int32_t(plan.modulation.size());
    v.filter_tracks = plan.filterTracks;
    v.frames = plan.frames;
    v.music_end = plan.musicEnd;
    v.limit = plan.limit;
    v.fade_frames = plan.fadeFrames;
    v.body_frames = plan.bodyFrames;
    v.onsets = plan.onsets.data();
    v.onset_count = int32_t(plan.onsets.size());
    v.volume_group = plan.volumeGroup;
}

} // namespace clay
