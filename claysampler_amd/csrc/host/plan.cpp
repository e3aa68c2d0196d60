// Job planner: one MIDI performance through one instrument, turned into the
// voice spans and control tracks the render core consumes.
//
// The planner is an event loop over the job's MIDI (frames are already
// integers). Between events it keeps three kinds of state:
//   - the channel: controller values, held keys, pedals, the bend in force;
//   - the voices sounding, each with the span it is extending;
//   - the plan being written: finished spans, control tracks, bindings.
// A span is cut wherever its parameters change mid-voice: at a bend change
// (the pitch ramp restarts) and where a sustain loop is left at release.
// Control tracks are piecewise-linear point lists with an optional smoothing
// filter, emitted in the form and with the timing rules of the reference
// (src/render/plan.cpp:114-216); voice parameters follow its trigger rules
// (:693-830) and release rules (:832-870). The features the instrument loader
// accepts are the ones planned here (see instrument.h).
#include "plan.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <map>
#include <random>
#include <stdexcept>
#include <tuple>

namespace clay {
namespace {

constexpr int kNoEnd = INT_MAX;
constexpr float kSilentLevel = 1e-4f;  // a release below this level has reached silence
constexpr int kBlock = 1024;           // sfizz_render's processing block: events ramp within it
constexpr float kMasterDb = -7.35f;    // output level of the orig/custom policies (sfizz's default volume)
constexpr float kQuarterPi = .7853981633974483f;
constexpr float kPanCompensation = 1.4125375446227544f;  // +3 dB of the stereo matrix
constexpr float kStereoGain = .99881488f;                // sfizz's stereo position gain at centre

// ------------------------------------------------------------ control tracks

// A controller change as sfizz_render applies it: a linear ramp from the
// start of the event's 1024-frame block (not before the previous event) to
// the event itself.
struct Ramp {
    int first, last;
    float from, to;
};

// Source 128 is the pitch bend; CC 121 resets a controller to its initial value.
template<class Map>
std::pair<float, std::vector<Ramp>> controllerRamps(const Midi& midi, int source, float initial, Map map, bool settle)
{
    float value = map(initial), start = value;
    std::vector<Ramp> ramps;
    int previousFrame = 0;
    for (const auto& e : midi.events) {
        float raw;
        if (e.kind == EventKind::control && e.a == 121)
            raw = initial;
        else if (source == 128 && e.kind == EventKind::bend)
            raw = float(e.a);
        else if (e.kind == EventKind::control && e.a == source)
            raw = float(e.b);
        else
            continue;
        const float next = map(raw);
        if (settle && e.frame == 0)
            start = next;  // the natural policy's leading defaults apply from frame 0
        else if (next != value)
            ramps.push_back({std::max(previousFrame, e.frame / kBlock * kBlock), e.frame, value, next});
        value = next;
        previousFrame = e.frame;
    }
    return {start, ramps};
}

// Appends a point unless the track already passes through it with that slope.
void appendPoint(std::vector<clay_control_point>& track, int frame, float value, float step)
{
    clay_control_point& last = track.back();
    if (last.frame == frame)
        last = {frame, value, step};
    else if (last.step != step || std::abs(last.value + (frame - last.frame) * step - value) > 1e-6f)
        track.push_back({frame, value, step});
}

// The track of a controller's ramps; with smoothing (sfizz's one-pole
// controller smoother, `smooth` frames) the target moves in 16-frame groups
// and the value follows it linearly.
std::vector<clay_control_point> trackOf(float start, const std::vector<Ramp>& ramps, int smooth)
{
    std::vector<clay_control_point> track{{0, start, 0}};
    if (smooth < 2) {
        for (const auto& r : ramps) {
            if (r.first < r.last)
                appendPoint(track, r.first, r.from, (r.to - r.from) / (r.last - r.first));
            appendPoint(track, r.last, r.to, 0);
        }
        return track;
    }
    constexpr int kGroup = 16;
    const int span = std::max(kGroup, smooth);
    float current = start, target = start, slope = 0;
    size_t next = 0;
    for (int frame = 0; next < ramps.size() || current != target; frame += kGroup) {
        if (current == target && next < ramps.size() && frame < ramps[next].first)
            frame = std::max(frame, ramps[next].first / kGroup * kGroup);  // idle until the next ramp
        // Where the controller stands at the end of this group.
        const int groupEnd = frame + kGroup - 1;
        float aim = target;
        while (next < ramps.size() && groupEnd >= ramps[next].last)
            aim = ramps[next++].to;
        if (next < ramps.size() && groupEnd >= ramps[next].first) {
            const Ramp& r = ramps[next];
            aim = r.from + (r.to - r.from) * (float(groupEnd - r.first) / (r.last - r.first));
        }
        if (aim != target) {
            target = aim;
            slope = (target - current) / span;
        }
        if (current == target || slope == 0) {
            appendPoint(track, frame, current, 0);
            continue;
        }
        appendPoint(track, frame, current + slope, slope);
        const float groups = std::abs((target - current) / slope);
        if (groups <= kGroup) {
            appendPoint(track, frame + std::max(0, int(std::ceil(groups)) - 1), target, 0);
            current = target;
        } else {
            current += kGroup * slope;
        }
    }
    return track;
}

// ------------------------------------------------------------ bends

struct BendChange {
    int begin = 0, end = 0, from = 0, to = 0;  // a ramp from `from` at begin to `to` at end
};

// Each bend value change ramps within its block, then holds (src/render/plan.cpp:114-134).
std::vector<BendChange> bendChanges(const Midi& midi)
{
    std::vector<BendChange> out;
    int previousFrame = 0, value = 0;
    for (const auto& e : midi.events) {
        const bool reset = e.kind == EventKind::control && e.a == 121;
        if (e.kind != EventKind::bend && !reset)
            continue;
        const int next = reset ? 0 : e.a;
        if (next != value) {
            out.push_back({std::max(previousFrame, e.frame / kBlock * kBlock), e.frame, value, next});
            out.push_back({e.frame, e.frame, next, next});
        }
        previousFrame = e.frame;
        value = next;
    }
    return out;
}

// ------------------------------------------------------------ envelopes and source travel
// Span and envelope-shape records share these field names (clay_voice_span,
// clay_envelope_shape), so one template serves both.

template<class E>
float levelAt(const E& e, int frame)
{
    if (e.release_start >= 0 && frame >= e.release_start) {
        if (frame >= e.fade_start)
            return std::max(0.f, e.fade_value - (frame - e.fade_start + 1) * e.fade_step);
        return e.release_value * std::exp((frame - e.release_start + 1) * e.release_rate);
    }
    if (frame < e.attack_start)
        return e.start;
    if (frame < e.hold_start)
        return std::min(1.f, e.start + (frame - e.attack_start + 1) * e.attack_step);
    if (frame < e.decay_start)
        return 1;
    if (e.decay_rate == 0)
        return e.sustain;
    return std::max(e.sustain, std::exp((frame - e.decay_start + 1) * e.decay_rate));
}

// Exponential release from `level` (-9 nepers over `seconds`), then a linear
// 50 ms fade from where it falls below the silent level.
template<class E>
void startReleaseAt(E& e, int frame, int rate, float level, float seconds)
{
    e.release_start = frame;
    e.release_value = level;
    e.release_rate = -9 / (std::clamp(seconds, .006f, 100.f) * rate);
    const int toSilence = level > kSilentLevel ? int(std::log(kSilentLevel / level) / e.release_rate) : 0;
    e.fade_start = frame + toSilence;
    e.fade_value = level * std::exp(e.release_rate * toSilence);
    e.fade_step = std::max(kSilentLevel, e.fade_value) / (rate * .05f);
}

// Source frames covered in `frames` output frames (the bend ramp is a
// geometric progression of the step).
double sourceTravel(const clay_voice_span& s, int frames)
{
    const double step = s.step * s.minimum_pitch_ratio;
    if (s.pitch_slope == 0)
        return step * frames;
    return step * std::expm1(s.pitch_slope * frames) / std::expm1(s.pitch_slope);
}

double insideLoop(const clay_voice_span& s, double position)
{
    if (s.loop_end > 0 && position >= s.loop_end)
        return s.loop_first + std::fmod(position - s.loop_first, s.loop_end - s.loop_first);
    return position;
}

// The frame where an unlooped sample runs out, or `envelopeEnd` if sooner
// (src/render/plan.cpp:264-280).
int spanEnd(const clay_voice_span& s, int envelopeEnd)
{
    if (s.loop_end > 0 || s.generator != CLAY_GEN_NONE)
        return envelopeEnd;
    const int rampFrames = std::max(0, s.bend_end - s.begin);
    const double rampTravel = sourceTravel(s, rampFrames);
    const double left = s.source_end - s.position;
    const double step = s.step * s.minimum_pitch_ratio;
    const double frames =
        rampFrames && s.pitch_slope != 0 && left <= rampTravel
            ? std::ceil(std::log1p(std::max(0., left) * std::expm1(s.pitch_slope) / step) / s.pitch_slope)
            : rampFrames + std::ceil((left - rampTravel) / (step * std::exp(s.pitch_slope * rampFrames)));
    if (frames > kNoEnd - s.begin)
        return envelopeEnd;
    return std::min(envelopeEnd, s.begin + std::max(0, int(frames)));
}

// Velocity to gain: squared velocity, or the region's curve, scaled by amp_veltrack.
float velocityGain(const Region& r, int velocity)
{
    const float v = float(velocity) / 127;
    float gain = v * v;
    if (!r.velocityCurve.empty()) {
        const auto hi = std::lower_bound(r.velocityCurve.begin(), r.velocityCurve.end(), velocity,
                                         [](const auto& p, int x) { return p.first < x; });
        if (hi->first == velocity)
            gain = hi->second;
        else {
            const auto lo = hi - 1;
            gain = lo->second + float(velocity - lo->first) / (hi->first - lo->first) * (hi->second - lo->second);
        }
    }
    const float reduction = std::abs(r.velocityTrack) * (1 - gain);
    return r.velocityTrack < 0 ? reduction : 1 - reduction;
}

// ------------------------------------------------------------ the planner

class JobPlanner {
public:
    JobPlanner(const Instrument& instrument, AudioBank& bank, int rate, unsigned seed, int limit, TailPolicy tail)
        : inst_(instrument), bank_(bank), rate_(rate), limit_(limit), natural_(tail == TailPolicy::natural),
          controllers_(instrument.ccDefault), noise_(seed ^ 0x85ebca6b)
    {
    }

    Plan run(const Midi& midi)
    {
        midi_ = &midi;
        anyBend_ = std::any_of(midi.events.begin(), midi.events.end(),
                               [](const Event& e) { return e.kind == EventKind::bend && e.a != 0; });
        // Tracks 0..2: volume (CC7), expression (CC11), pan (CC10), constant
        // when the instrument binds that controller itself.
        for (int cc : {7, 11, 10}) {
            if (inst_.ccUsed[size_t(cc)])
                plan_.controls.push_back({{0, cc == 10 ? 0.f : 1.f, 0}});
            else
                plan_.controls.push_back(controllerTrack(cc, cc == 10 ? 1 : 4, rate_ / 100));
        }
        const auto bends = bendChanges(midi);
        size_t bend = 0;
        for (const auto& e : midi.events) {
            for (; bend < bends.size() && bends[bend].begin <= e.frame; ++bend)
                applyBendChange(bends[bend]);
            retireEnded(e.frame);
            if (e.kind == EventKind::on)
                noteOn(e);
            else if (e.kind == EventKind::off)
                noteOff(e);
            else if (e.kind == EventKind::control)
                controller(e);
        }
        retireEnded(midi.end);
        for (auto& v : voices_) {
            // The natural policy's MIDI ends its notes itself (pedals lifted at the end).
            if (!natural_)
                release(v, midi.end);
            if (v.span.end == kNoEnd && limit_ == kNoEnd)
                throw std::runtime_error("a looping one-shot or release sample has no finite end");
            finish(v.span, v.span.end);
        }
        if (!natural_)
            plan_.frames = std::max(plan_.frames, midi.end);
        if (plan_.spans.empty())
            throw std::runtime_error("no playable notes for this SFZ and MIDI");
        return std::move(plan_);
    }

private:
    struct Voice {
        clay_voice_span span;
        const Region* region;
        int key, velocity;
        double baseStep;        // source frames per output frame without bend
        float releaseSeconds;
        bool keyReleased = false;  // note-off seen (the pedals may still hold it)
        bool sostenuto = false;    // held by the sostenuto pedal
    };

    // ---------------------------------------------------- plan output

    // Writes a span (cut at `end` and at the render limit) to the plan.
    void finish(clay_voice_span s, int end)
    {
        s.end = std::min(s.end, end);
        plan_.tailClipped |= s.end > limit_ && s.begin < s.end;
        s.end = std::min(s.end, limit_);
        if (s.begin >= s.end)
            return;
        plan_.frames = std::max(plan_.frames, s.end);
        plan_.spans.push_back(s);
    }

    // Voices whose spans have ended by `frame` leave the pool, in order.
    void retireEnded(int frame)
    {
        auto keep = voices_.begin();
        for (auto& v : voices_) {
            if (v.span.end > frame)
                *keep++ = v;
            else
                finish(v.span, v.span.end);
        }
        voices_.erase(keep, voices_.end());
    }

    // Ends the voice's current span at `frame` and continues it from there.
    void cut(Voice& v, int frame, double position)
    {
        finish(v.span, frame);
        v.span.begin = frame;
        v.span.position = position;
    }

    std::vector<clay_control_point> controllerTrack(int cc, int curve, int smooth)
    {
        const auto [start, ramps] = controllerRamps(*midi_, cc, inst_.ccDefault[size_t(cc)],
                                                    [curve](float x) { return curveValue(curve, x); }, natural_);
        return trackOf(start, ramps, smooth);
    }

    // Index of the plan track for a controller binding (shared by equal bindings).
    int bindingTrack(const Binding& b)
    {
        const auto key = std::make_tuple(b.source, b.curve, b.smoothMs);
        const auto found = bindingTracks_.find(key);
        if (found != bindingTracks_.end())
            return found->second;
        const int index = int(plan_.controls.size());
        bindingTracks_.emplace(key, index);
        plan_.controls.push_back(controllerTrack(b.source, b.curve, b.smoothMs * rate_ / 1000));
        return index;
    }

    // Pitch track (cents) of a smoothed bend (bend_smooth).
    int smoothBendTrack(const Region& r)
    {
        const auto key = std::make_tuple(r.bendUp, r.bendDown, r.bendSmoothMs);
        const auto found = bendTracks_.find(key);
        if (found != bendTracks_.end())
            return found->second;
        const int index = int(plan_.controls.size());
        bendTracks_.emplace(key, index);
        const auto cents = [&r](float value) {
            const float bend = std::max(-8191.f, value) / 8191;
            return bend >= 0 ? bend * r.bendUp : -bend * r.bendDown;
        };
        const auto [start, ramps] = controllerRamps(*midi_, 128, 0, cents, natural_);
        plan_.controls.push_back(trackOf(start, ramps, r.bendSmoothMs * rate_ / 1000));
        return index;
    }

    // ---------------------------------------------------- controllers and parameters

    // A binding's source value through its curve (the key steps 140/141 unshaped).
    float sourceValue(const Binding& b, int key, int velocity) const
    {
        if (b.source == 140)
            return keyStep_;
        if (b.source == 141)
            return std::abs(keyStep_);
        float raw = 0;
        if (b.source < 128)
            raw = controllers_[size_t(b.source)];
        else if (b.source == 133)
            raw = float(key);
        else if (b.source == 131)
            raw = float(velocity);  // note-on velocity (every voice here starts on a note)
        else if (b.source == 134)
            raw = keysHeld_ > 0 ? 127.f : 0.f;
        else if (const auto d = inst_.extendedDefaults.find(b.source); d != inst_.extendedDefaults.end())
            raw = d->second;
        return curveValue(b.curve, raw);
    }

    // A voice parameter at note-on: base value, amplitude-envelope velocity
    // offset, plus every binding to it.
    float parameter(const Region& r, ModTarget target, float base, int key, int velocity) const
    {
        if (target >= ModTarget::envDelay && target <= ModTarget::envRelease)
            base += r.ampVelocity[size_t(int(target) - int(ModTarget::envDelay))] * (velocity / 127.f);
        for (const auto& b : r.bindings)
            if (b.target == target)
                base += b.depth * sourceValue(b, key, velocity);
        return base;
    }

    bool pedalDown(const Region& r) const
    {
        return r.sustainSwitch && controllers_[size_t(r.sustainCC)] >= r.sustainLow;
    }

    bool selected(const Region& r, int velocity) const
    {
        if (velocity < r.velocityLow || velocity > r.velocityHigh)
            return false;
        for (const auto& w : r.conditions)
            if (controllers_[size_t(w.cc)] < w.low || controllers_[size_t(w.cc)] > w.high)
                return false;
        return true;
    }

    // ---------------------------------------------------- pitch

    // Step and pitch ramp of the voice's current span under the bend in force.
    void bendSpan(Voice& v) const
    {
        clay_voice_span& s = v.span;
        const Region& r = *v.region;
        s.pitch_slope = 0;
        if (r.bendSmoothMs > 0) {  // the bend reaches the voice through its smoothed pitch track
            s.bend_end = s.begin;
            s.step = v.baseStep;
            return;
        }
        const auto cents = [&r](int value) {
            const double bend = std::max(-8191, value) / 8191.0;
            return bend >= 0 ? bend * r.bendUp : -bend * r.bendDown;
        };
        const double to = cents(bend_.to);
        s.bend_end = bend_.end;
        if (s.begin >= bend_.end) {
            s.step = v.baseStep * std::exp2(to / 1200);
            return;
        }
        const double from = cents(bend_.from);
        const double slope = (to - from) / (bend_.end - bend_.begin);  // cents per frame
        s.pitch_slope = slope * (std::log(2.) / 1200);
        s.step = v.baseStep * std::exp2((from + (s.begin - bend_.begin) * slope) / 1200);
    }

    int fadeEnd(const clay_voice_span& s) const { return s.release_start < 0 ? kNoEnd : s.fade_start + rate_ / 20 + 1; }

    void applyBendChange(const BendChange& change)
    {
        retireEnded(change.begin);
        bend_ = change;
        for (auto& v : voices_) {
            if (v.region->bendSmoothMs > 0)
                continue;
            const int from = std::max(change.begin, v.span.begin);
            cut(v, from, v.span.position + sourceTravel(v.span, from - v.span.begin));
            bendSpan(v);
            v.span.end = spanEnd(v.span, fadeEnd(v.span));
        }
    }

    // ---------------------------------------------------- starting voices

    // Filter and pitch envelope shapes (fileg_, pitcheg_) of a new voice.
    void modEnvelopes(Voice& v) const
    {
        const Region& r = *v.region;
        for (int slot = 0; slot < 2; ++slot) {
            const ModEnvelope& src = r.modEnvelopes[size_t(slot)];
            if (!src.used)
                continue;
            const int first = int(slot == 0 ? ModTarget::filterEnvDelay : ModTarget::pitchEnvDelay);
            Stages st;
            for (int k = 0; k < stageCount; ++k) {
                const float base = src.stages[size_t(k)] + src.velocity[size_t(k)] * (v.velocity / 127.f);
                const float value = parameter(r, ModTarget(first + k), base, v.key, v.velocity);
                st[size_t(k)] = std::clamp(value, 0.f, k == stageSustain || k == stageStart ? 1.f : 100.f);
            }
            clay_envelope_shape& e = v.span.extra_envelope[slot];
            v.span.active_envelope[slot] = 1;
            e.start = st[stageStart];
            e.attack_start = v.span.begin + int(st[stageDelay] * rate_);
            const float attack = st[stageAttack] * rate_;
            e.attack_step = attack > 0 ? 1 / attack : 1;
            e.hold_start = e.attack_start + std::max(0, int(std::ceil((1 - e.start) * attack)) - 1);
            e.decay_start = e.hold_start + int(st[stageHold] * rate_);
            e.decay_rate = st[stageDecay] > 0 ? -9 / (std::max(.006f, st[stageDecay]) * rate_) : 0;
            e.sustain = st[stageSustain];
            e.release_start = -1;
            e.release_seconds = st[stageRelease];
            e.depth = src.depth + src.velocityDepth * (v.velocity / 127.f);
        }
    }

    int newFilterTrack(clay_voice_span& s)
    {
        if (s.filter_track < 0)
            s.filter_track = plan_.filterTracks++;
        return s.filter_track;
    }

    void startVoice(const Region& r, int frame, int key, int velocity)
    {
        const Sample* sample = r.generator == Generator::none ? &bank_.decode(r.sample) : nullptr;
        const auto param = [&](ModTarget t, float base) { return parameter(r, t, base, key, velocity); };
        const auto seconds = [&](ModTarget t, float base) { return std::clamp(param(t, base), 0.f, 100.f); };
        Voice v{};
        v.region = &r;
        v.key = key;
        v.velocity = velocity;
        v.releaseSeconds = seconds(ModTarget::envRelease, r.amp[stageRelease]);
        const double rootRate = sample ? sample->rate : 440 * std::exp2((r.keyRoot - 69) / 12.);
        const float cents = r.keyTrack * (key - r.keyRoot) + r.tune + 100 * r.transpose;
        v.baseStep = (rootRate / rate_) * std::exp2(cents / 1200.0);

        clay_voice_span& s = v.span;
        s.sample = sample ? sample->offset : 0;
        s.channels = sample ? sample->channels : 1;
        s.generator = int32_t(r.generator);
        if (r.generator == Generator::noise) {
            s.noise_seed = noise_();
            s.noise_start = frame;
        }
        s.begin = frame + int(seconds(ModTarget::delay, r.delay) * rate_);
        s.source_end = r.end;
        if (sample) {
            const double offset = param(ModTarget::offset, float(r.offset));
            s.position = std::trunc(std::clamp(offset, 0., double(r.end)));
        } else {
            s.position = r.oscillatorPhase - std::floor(r.oscillatorPhase);
        }
        bendSpan(v);
        if (sample && (r.loop == LoopMode::continuous || r.loop == LoopMode::sustain)) {
            s.loop_first = r.loopStart;
            s.loop_end = r.loopEnd;
            s.loop_crossfade = int(std::lround(std::min<double>(r.loopStart, double(r.loopCrossfade) * sample->rate)));
        }
        // Amplitude envelope.
        s.start = std::clamp(param(ModTarget::envStart, r.amp[stageStart]), 0.f, 1.f);
        s.attack_start = s.begin + int(seconds(ModTarget::envDelay, r.amp[stageDelay]) * rate_);
        const float attack = seconds(ModTarget::envAttack, r.amp[stageAttack]) * rate_;
        s.attack_step = attack > 0 ? 1 / attack : 1;
        s.hold_start = s.attack_start + std::max(0, int(std::ceil((1 - s.start) * attack)) - 1);
        s.decay_start = s.hold_start + int(seconds(ModTarget::envHold, r.amp[stageHold]) * rate_);
        const float decay = seconds(ModTarget::envDecay, r.amp[stageDecay]);
        s.decay_rate = decay > 0 ? -9 / (std::max(.006f, decay) * rate_) : 0;
        s.sustain = std::clamp(param(ModTarget::envSustain, r.amp[stageSustain]), 0.f, 1.f);
        s.release_start = -1;
        modEnvelopes(v);
        // Filters and EQ bands.
        s.filter_track = -1;
        s.sample_rate = rate_;
        for (int f = 0; f < 2; ++f) {
            const FilterSettings& fs = r.filters[size_t(f)];
            if (fs.type == FilterType::none)
                continue;
            newFilterTrack(s);
            const float tracking = fs.keyTrack * (key - fs.keyCenter) +
                                   param(f == 0 ? ModTarget::filterVelocity : ModTarget::filterVelocity2,
                                         fs.velocityTrack) * velocity / 127.f;
            const float cutoff = fs.cutoff * std::exp2(std::clamp(tracking, -24000.f, 24000.f) / 1200);
            s.filters[f] = {int32_t(fs.type), std::clamp(cutoff, 0.f, 20000.f), fs.resonance, 0, 0};
        }
        static constexpr float kEqCentres[3] = {50, 500, 5000};
        for (int band = 0; band < 3; ++band) {
            const EqSettings& eq = r.eq[size_t(band)];
            if (!eq.used)
                continue;
            newFilterTrack(s);
            s.filters[band + 2] = {int32_t(FilterType::peak), eq.frequency == 0 ? kEqCentres[band] : eq.frequency, 0,
                                   eq.bandwidth, eq.gain};
        }
        // Bindings that modulate the voice while it sounds; a pitch modulation
        // range makes the voice a variable-pitch chain.
        s.first_modulation = int(plan_.modulation.size());
        bool stereoModulated = false;
        float pitchRange = std::abs(s.extra_envelope[1].depth);
        for (const auto& b : r.bindings) {
            const bool atNoteOn = (b.target >= ModTarget::envDelay && b.target < ModTarget::cutoff) ||
                                  (b.target >= ModTarget::filterVelocity && b.target < ModTarget::filterEnvDepth);
            if (atNoteOn)
                continue;
            stereoModulated |= b.target == ModTarget::width || b.target == ModTarget::position;
            if (b.target == ModTarget::pitch || b.target == ModTarget::pitchEnvDepth)
                pitchRange += std::abs(b.source < 128 ? b.depth : b.depth * sourceValue(b, key, velocity));
            if (b.source < 128)
                plan_.modulation.push_back({bindingTrack(b), int32_t(b.target), b.depth, -1});
            else
                plan_.modulation.push_back({-1, int32_t(b.target), b.depth * sourceValue(b, key, velocity), -1});
        }
        if (r.bendSmoothMs > 0 && anyBend_) {
            pitchRange += std::max(std::abs(r.bendUp), std::abs(r.bendDown));
            plan_.modulation.push_back({smoothBendTrack(r), CLAY_MOD_PITCH, 1, -1});
        }
        s.minimum_pitch_ratio = std::exp2(-std::min(24000.f, pitchRange) / 1200);
        if (pitchRange > 0) {
            s.variable_pitch = 1;
            newFilterTrack(s);
        }
        s.modulation_count = int(plan_.modulation.size()) - s.first_modulation;
        // Level and stereo placement.
        const float db = (natural_ ? 0.f : kMasterDb) + r.volumeDb + 3;
        const float gain = r.amplitude * velocityGain(r, velocity) * std::pow(10.f, db / 20);
        const bool matrix = s.channels == 2 && (stereoModulated || r.width != 1 || r.position != 0);
        s.stereo_matrix = matrix;
        s.gain = gain * (matrix ? kPanCompensation : s.channels == 1 ? 1.f : kStereoGain);
        s.pan = r.pan;
        s.width = r.width;
        s.stereo_position = r.position;
        s.left = s.gain * std::cos((1 + r.pan) * kQuarterPi);
        s.right = s.gain * std::cos((1 - r.pan) * kQuarterPi);
        s.width_direct = matrix ? std::cos((1 - r.width) * kQuarterPi) : 1.f;
        s.width_cross = matrix ? std::cos((1 + r.width) * kQuarterPi) : 0.f;
        s.position_left = matrix ? std::cos((1 + r.position) * kQuarterPi) : 1.f;
        s.position_right = matrix ? std::cos((1 - r.position) * kQuarterPi) : 1.f;
        s.end = spanEnd(s, kNoEnd);
        voices_.push_back(v);
    }

    // ---------------------------------------------------- releasing voices

    void release(Voice& v, int frame)
    {
        if (v.span.release_start >= 0 || v.region->loop == LoopMode::oneShot)
            return;
        clay_voice_span& s = v.span;
        if (frame >= s.end)
            return;
        if (frame < s.begin) {  // released during its delay: never sounds
            s.end = s.begin;
            return;
        }
        const float level = levelAt(s, frame - 1);
        if (v.region->loop == LoopMode::sustain && s.loop_end > 0) {
            // Leaves the sustain loop: the source runs on from the loop position.
            cut(v, frame, insideLoop(s, s.position + sourceTravel(s, frame - s.begin)));
            bendSpan(v);
            s.loop_end = 0;
        }
        startReleaseAt(s, frame, rate_, level, v.releaseSeconds);
        for (int slot = 0; slot < 2; ++slot) {
            clay_envelope_shape& e = s.extra_envelope[slot];
            if (s.active_envelope[slot] && e.release_start < 0)
                startReleaseAt(e, frame, rate_, levelAt(e, frame - 1), e.release_seconds);
        }
        s.end = spanEnd(s, fadeEnd(s));
    }

    // Voices whose key is up and that no pedal holds any more.
    void releaseUnheld(int frame, int cc, float before)
    {
        const bool sostenutoLifted = natural_ && cc == 66 && !sostenutoDown_;
        for (auto& v : voices_) {
            const Region& r = *v.region;
            const bool sustainLifted = r.sustainSwitch && r.sustainCC == cc && before >= r.sustainLow;
            if (v.keyReleased && (sustainLifted || sostenutoLifted) && !pedalDown(r) && !v.sostenuto)
                release(v, frame);
        }
    }

    void setSostenuto(bool down)
    {
        if (sostenutoDown_ == down)
            return;
        sostenutoDown_ = down;
        for (auto& v : voices_)
            v.sostenuto = down && !v.keyReleased;
    }

    // ---------------------------------------------------- events

    void noteOn(const Event& e)
    {
        if (!keyDown_[size_t(e.a)])
            ++keysHeld_;
        keyStep_ = lastKey_ < 0 ? 0.f : float(e.a - lastKey_);
        lastKey_ = e.a;
        keyDown_[size_t(e.a)] = true;
        for (int index : inst_.regionsOfKey[size_t(e.a)]) {
            const Region& r = inst_.regions[size_t(index)];
            if (selected(r, e.b))
                startVoice(r, e.frame, e.a, e.b);
        }
    }

    void noteOff(const Event& e)
    {
        if (!keyDown_[size_t(e.a)])
            return;
        keyDown_[size_t(e.a)] = false;
        --keysHeld_;
        for (auto& v : voices_)
            if (v.key == e.a) {
                v.keyReleased = true;
                if (!pedalDown(*v.region) && !v.sostenuto)
                    release(v, e.frame);
            }
    }

    void controller(const Event& e)
    {
        if (e.a == 121) {  // reset all controllers
            const float sostenutoBefore = controllers_[66];
            controllers_ = inst_.ccDefault;
            if (natural_) {
                setSostenuto(false);
                releaseUnheld(e.frame, 66, sostenutoBefore);
            }
            for (auto& v : voices_)
                if (v.keyReleased && !pedalDown(*v.region))
                    release(v, e.frame);
            return;
        }
        if (!natural_ && e.a == 66 && e.b != 0)
            throw std::runtime_error("unsupported MIDI sostenuto CC66");
        if (natural_ && (e.a == 120 || e.a == 123)) {  // all sound / all notes off: everything stops
            for (auto& v : voices_)
                finish(v.span, e.frame);
            voices_.clear();
            keyDown_.fill(false);
            keysHeld_ = 0;
            return;
        }
        if (e.a >= 120)
            throw std::runtime_error("unsupported MIDI channel mode CC" + std::to_string(e.a));
        const bool sostenuto = natural_ && e.a == 66;
        if (!inst_.ccUsed[size_t(e.a)] && !sostenuto)
            return;
        const float before = controllers_[size_t(e.a)];
        controllers_[size_t(e.a)] = float(e.b);
        if (sostenuto)
            setSostenuto(e.b == 127);
        releaseUnheld(e.frame, e.a, before);
    }

    const Instrument& inst_;
    AudioBank& bank_;
    const int rate_, limit_;
    const bool natural_;
    const Midi* midi_ = nullptr;
    bool anyBend_ = false;
    // Channel.
    std::array<float, 128> controllers_;
    std::array<bool, 128> keyDown_{};
    int keysHeld_ = 0, lastKey_ = -1;
    float keyStep_ = 0;
    bool sostenutoDown_ = false;
    BendChange bend_;
    std::mt19937 noise_;  // noise generator seeds
    // Voices and output.
    std::vector<Voice> voices_;
    std::map<std::tuple<int, int, int>, int> bindingTracks_;
    std::map<std::tuple<float, float, int>, int> bendTracks_;
    Plan plan_;
};

} // namespace

Plan planRender(const Instrument& instrument, AudioBank& bank, const Midi& midi, int rate, unsigned seed, int limit,
                TailPolicy tail)
{
    return JobPlanner(instrument, bank, rate, seed, limit, tail).run(midi);
}

void makeView(const Plan& plan, PlanView& out)
{
    out.points.clear();
    out.offsets.assign(1, 0);
    for (const auto& track : plan.controls) {
        out.points.insert(out.points.end(), track.begin(), track.end());
        out.offsets.push_back(int32_t(out.points.size()));
    }
    clay_plan_view& v = out.view;
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
