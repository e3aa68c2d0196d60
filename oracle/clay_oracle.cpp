// FP64 scalar oracle of the ClaySampler render core -- test infrastructure,
// never part of the product. Each function restates one piece of the
// reference kernel math in double precision; the file:line it follows is
// given at each function. Discrete decisions (phase index, loop wrap,
// envelope stage, tail thresholds) use the same rules as the reference.
#include "clay_oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

thread_local std::string g_error;

constexpr int kPhases = CLAY_SINC_PHASES;
constexpr int kTailBlock = 1024;  // src/cuda/render_types.h:92
constexpr double kPi = 3.14159265358979323846;

struct Stereo {
    double l = 0, r = 0;
};

// src/cuda/renderer.cpp:347-376, generalised: centre tap taps/2-1, half width taps/2.
std::vector<float> sincTable(int taps)
{
    const double beta = 8.5;
    const auto bessel = [](double v) {
        double sum = 1, term = 1;
        for (int i = 1; i < 30; ++i) {
            term *= v * v / (4 * i * i);
            sum += term;
            if (term < 1e-14)
                break;
        }
        return sum;
    };
    const double norm = bessel(beta);
    const int centre = taps / 2 - 1;
    const double half = taps / 2;
    std::vector<float> table(size_t(taps) * (kPhases + 1));
    for (int t = 0; t < taps; ++t)
        for (int p = 0; p <= kPhases; ++p) {
            const double x = t - centre - double(p) / kPhases;
            const double f = x / half;
            const double window = bessel(beta * std::sqrt(1 - f * f)) / norm;
            table[size_t(t) * (kPhases + 1) + size_t(p)] = float(window * (x == 0 ? 1 : std::sin(kPi * x) / (kPi * x)));
        }
    return table;
}

// Resolved plan of one job: control tracks as absolute point ranges.
struct Job {
    const clay_plan_view* plan;
    std::vector<int> trackFirst, trackEnd;

    explicit Job(const clay_plan_view* p) : plan(p)
    {
        for (int t = 0; t < p->track_count; ++t) {
            trackFirst.push_back(p->track_offsets[t]);
            trackEnd.push_back(p->track_offsets[t + 1]);
        }
    }

    // src/cuda/render.cu:107-118
    double controlAt(int track, int frame) const
    {
        int lo = trackFirst[size_t(track)], hi = trackEnd[size_t(track)];
        while (lo + 1 < hi) {
            const int mid = lo + (hi - lo) / 2;
            if (plan->points[mid].frame <= frame)
                lo = mid;
            else
                hi = mid;
        }
        const auto& p = plan->points[lo];
        return double(p.step) * double(frame - p.frame) + double(p.value);
    }

    // src/cuda/render.cu:120-126
    double modulationAt(const clay_modulation& m, int frame) const
    {
        double v = m.track < 0 ? double(m.depth) : double(m.depth) * controlAt(m.track, frame);
        if (m.shape_track >= 0)
            v *= controlAt(m.shape_track, frame);
        return v;
    }

    // src/cuda/render.cu:189-193
    double jobGain(int frame) const { return controlAt(0, frame) * controlAt(1, frame); }
    double panBase(int frame) const { return controlAt(2, frame); }
};

// src/cuda/render.cu:6-19 (amplitude envelope) in double.
template<class E>
double envelope(const E& v, int frame)
{
    if (v.release_start >= 0 && frame >= v.release_start) {
        if (frame >= v.fade_start)
            return std::max(0.0, double(v.fade_value) - double(frame - v.fade_start + 1) * double(v.fade_step));
        return double(v.release_value) * std::exp(double(frame - v.release_start + 1) * double(v.release_rate));
    }
    if (frame < v.attack_start)
        return v.start;
    if (frame < v.hold_start)
        return std::min(1.0, double(v.start) + double(frame - v.attack_start + 1) * double(v.attack_step));
    if (frame < v.decay_start)
        return 1;
    if (v.decay_rate == 0)
        return v.sustain;
    return std::max(double(v.sustain), std::exp(double(frame - v.decay_start + 1) * double(v.decay_rate)));
}

struct Renderer {
    const int16_t* bank;
    uint64_t bankCount;
    int taps;
    std::vector<float> table;
    float silence = 0;  // voice-silence policy level (linear), 0 = off

    // src/cuda/render.cu:21-28
    Stereo fetch(const clay_voice_span& v, long long frame) const
    {
        if (frame < 0 || frame >= v.source_end)
            return {};
        const uint64_t i = v.sample + uint64_t(frame) * uint64_t(v.channels);
        const double left = bank[i] / 32768.0;
        return {left, v.channels == 1 ? left : bank[i + 1] / 32768.0};
    }

    // src/cuda/render.cu:30-61: nearest of 2048 phases, exact sample at integer positions.
    Stereo interpolateAt(const clay_voice_span& v, double pos) const
    {
        const double centreD = std::floor(pos);
        const long long centre = (long long)centreD;
        if (pos == centreD)
            return fetch(v, centre);
        const int phase = int(std::nearbyint((pos - centreD) * kPhases));
        const int first = taps / 2 - 1;
        Stereo out;
        for (int t = 0; t < taps; ++t) {
            const double w = table[size_t(t) * (kPhases + 1) + size_t(phase)];
            const Stereo s = fetch(v, centre - first + t);
            out.l += w * s.l;
            out.r += w * s.r;
        }
        return out;
    }

    // src/cuda/render.cu:63-100
    Stereo interpolate(const clay_voice_span& v, int frame, double pos) const
    {
        if (pos < 0) {
            const int elapsed = frame - v.begin;
            double distance = elapsed;
            if (v.pitch_slope != 0)
                distance = std::expm1(v.pitch_slope * elapsed) / std::expm1(v.pitch_slope);
            pos = v.position + distance * v.step;
        }
        if (v.generator != CLAY_GEN_NONE) {
            if (v.generator == CLAY_GEN_NOISE) {
                unsigned bits = v.noise_seed + unsigned(frame - v.noise_start);
                bits = (bits ^ (bits >> 16)) * 0x7feb352du;
                bits = (bits ^ (bits >> 15)) * 0x846ca68bu;
                bits ^= bits >> 16;
                const double x = double(bits >> 8) * (2.0 / 16777216.0) - 1;
                return {x, x};
            }
            const double x = v.generator == CLAY_GEN_SINE ? -std::sin(2 * kPi * (pos - std::floor(pos))) : 0.0;
            return {x, x};
        }
        const int period = v.loop_end - v.loop_first;
        if (v.loop_end > 0 && pos >= v.loop_end)
            pos = v.loop_first + std::fmod(pos - v.loop_first, double(period));
        if (pos >= v.source_end)
            return {};
        Stereo s = interpolateAt(v, pos);
        const int fadeStart = v.loop_end - v.loop_crossfade;
        if (v.loop_end > 0 && v.loop_crossfade > 0 && pos >= fadeStart) {
            const Stereo entry = interpolateAt(v, pos - period);
            const double phase = (pos - fadeStart) / v.loop_crossfade;
            const double blend = .5 - .5 * std::cos(kPi * phase);
            s.l += blend * (entry.l - s.l);
            s.r += blend * (entry.r - s.r);
        }
        return s;
    }

    // src/cuda/render.cu:130-136
    static Stereo panGains(const clay_voice_span& v, double offset)
    {
        if (offset == 0)
            return {v.left, v.right};
        const double p = std::min(1.0, std::max(-1.0, double(v.pan) + offset));
        return {double(v.gain) * std::cos((1 + p) * kPi / 4), double(v.gain) * std::cos((1 - p) * kPi / 4)};
    }

    // src/cuda/render.cu:138-187
    Stereo voiceAudio(const Job& job, const clay_voice_span& v, double pan, int frame, bool filterInput, double phase) const
    {
        double level = envelope(v, frame);
        if (level <= 0)
            return {};
        double db = 0, widthOffset = 0, positionOffset = 0;
        for (int j = v.first_modulation; j < v.first_modulation + v.modulation_count; ++j) {
            const auto& m = job.plan->modulation[j];
            if (m.target > CLAY_MOD_POSITION)
                continue;
            const double x = job.modulationAt(m, frame);
            switch (m.target) {
            case CLAY_MOD_AMPLITUDE: level *= x; break;
            case CLAY_MOD_VOLUME: db += x; break;
            case CLAY_MOD_PAN: pan += x; break;
            case CLAY_MOD_WIDTH: widthOffset += x; break;
            case CLAY_MOD_POSITION: positionOffset += x; break;
            default: break;
            }
        }
        if (db != 0)
            level *= std::pow(10.0, db / 20);
        const Stereo s = interpolate(v, frame, phase);
        if (filterInput && v.channels == 1)
            return {s.l * level, s.l * level};
        const Stereo g = panGains(v, pan);
        const double l = s.l * g.l, r = s.r * g.r;
        if (v.stereo_matrix) {
            double direct = v.width_direct, cross = v.width_cross;
            if (widthOffset != 0) {
                const double w = std::min(1.0, std::max(-1.0, double(v.width) + widthOffset));
                direct = std::cos((1 - w) * kPi / 4);
                cross = std::cos((1 + w) * kPi / 4);
            }
            double pl = v.position_left, pr = v.position_right;
            if (positionOffset != 0) {
                const double p = std::min(1.0, std::max(-1.0, double(v.stereo_position) + positionOffset));
                pl = std::cos((1 + p) * kPi / 4);
                pr = std::cos((1 - p) * kPi / 4);
            }
            return {(l * direct + r * cross) * (level * pl), (r * direct + l * cross) * (level * pr)};
        }
        return {l * level, r * level};
    }
};

struct Coefficients {
    double b0 = 0, b1 = 0, b2 = 0, a1 = 0, a2 = 0;
};

// src/cuda/render.cu:278-310 (RBJ-style biquads, 1-pole for low1/high1).
Coefficients coefficients(int type, double cutoff, double resonance, int rate, double bandwidth, double gain)
{
    const double w = 2 * kPi * std::min(std::min(20000.0, rate * .499), std::max(1.0, cutoff)) / rate;
    if (type == CLAY_FILTER_LOW1 || type == CLAY_FILTER_HIGH1) {
        const double p = std::exp(-w);
        if (type == CLAY_FILTER_LOW1)
            return {1 - p, 0, 0, -p, 0};
        const double h = (1 + p) * .5;
        return {h, -h, 0, -p, 0};
    }
    const double s = std::sin(w), c = std::cos(w);
    if (type == CLAY_FILTER_PEAK) {
        const double a = std::pow(10.0, std::min(60.0, std::max(-96.0, gain)) / 40);
        const double width = std::min(12.0, std::max(.01, bandwidth));
        const double alpha = s * std::sinh(std::min(20.0, std::log(2.0) / 2 * width * w / s));
        const double inv = 1 / (1 + alpha / a);
        const double mid = -2 * c * inv;
        return {(1 + alpha * a) * inv, mid, (1 - alpha * a) * inv, mid, (1 - alpha / a) * inv};
    }
    const double alpha = .5 * s * std::pow(10.0, -std::min(60.0, std::max(-60.0, resonance)) / 20);
    const double inv = 1 / (1 + alpha);
    const double a1 = -2 * c * inv, a2 = (1 - alpha) * inv;
    if (type == CLAY_FILTER_BAND2 || type == CLAY_FILTER_BAND4 || type == CLAY_FILTER_BAND6)
        return {alpha * inv, 0, -alpha * inv, a1, a2};
    if (type == CLAY_FILTER_NOTCH2)
        return {inv, a1, inv, a1, a2};
    const bool high = type == CLAY_FILTER_HIGH2 || type == CLAY_FILTER_HIGH4 || type == CLAY_FILTER_HIGH6;
    const double b = (1 + (high ? c : -c)) * .5 * inv;
    return {b, (high ? -2 : 2) * b, b, a1, a2};
}

// src/cuda/render.cu:312-319
int sections(int type)
{
    if (type == CLAY_FILTER_NONE) return 0;
    if (type == CLAY_FILTER_PEAK) return 1;
    if (type >= CLAY_FILTER_LOW6) return 3;
    if (type >= CLAY_FILTER_LOW4) return 2;
    return 1;
}

// Direct form II transposed section (src/cuda/render.cu:330-336).
double filterSample(double x, const Coefficients& c, double& z1, double& z2)
{
    const double y = c.b0 * x + z1;
    z1 = -c.a1 * y + (c.b1 * x + z2);
    z2 = -c.a2 * y + c.b2 * x;
    return y;
}

// All spans with the same filter track, in plan order (the reference links them
// with nextFilteredSpan: src/cuda/renderer.cpp:428-456).
std::vector<std::vector<int>> chains(const clay_plan_view& p)
{
    std::vector<std::vector<int>> out(size_t(std::max(0, p.filter_tracks)));
    for (int i = 0; i < p.span_count; ++i)
        if (p.spans[i].filter_track >= 0)
            out[size_t(p.spans[i].filter_track)].push_back(i);
    return out;
}

constexpr int kSilentChunks = CLAY_SILENT_CHUNKS;

// Voice-silence policy of one chain (no reference counterpart): chunks are the
// absolute 64-frame blocks the chain covers; a chunk is silent when every
// rendered output sample of the chain in it (float, non-finite = loud) is below
// the level and every one of its frames is at or after its span's release.
// The chain retires at the end of its first run of kSilentChunks silent chunks.
int chainSilence(const clay_plan_view& p, const std::vector<int>& chain,
                 const std::vector<std::vector<Stereo>>& filtered, float level)
{
    int end = 0;
    for (int s : chain)
        end = std::max(end, p.spans[s].end);
    if (!(level > 0))
        return end;
    int open = -1, silent = 0;
    float peak = 0;
    bool eligible = true;
    for (int s : chain) {
        const auto& v = p.spans[s];
        for (int f = v.begin; f < v.end; ++f) {
            if ((f >> 6) != open) {
                open = f >> 6;
                peak = 0;
                eligible = true;
            }
            const Stereo& x = filtered[size_t(s)][size_t(f - v.begin)];
            float a = std::max(std::fabs(float(x.l)), v.channels == 2 ? std::fabs(float(x.r)) : 0.f);
            if (!(a <= std::numeric_limits<float>::max()))
                a = std::numeric_limits<float>::infinity();
            peak = std::max(peak, a);
            eligible = eligible && v.release_start >= 0 && f >= v.release_start;
            if ((f & 63) == 63) {
                silent = eligible && peak < level ? silent + 1 : 0;
                open = -1;
                if (silent == kSilentChunks)
                    return f + 1;
            }
        }
    }
    return end;
}

// Mix of one job (src/cuda/render.cu:195-492 with one segment covering the job).
void mixJob(const Renderer& R, const Job& job, std::vector<Stereo>& mix, std::vector<int>* tails = nullptr)
{
    const auto& p = *job.plan;
    const int frames = p.frames;
    std::vector<Stereo> dry(static_cast<size_t>(frames)), wet(static_cast<size_t>(frames));
    // Variable pitch phase per chain span (src/cuda/render.cu:222-266).
    std::vector<std::vector<double>> phase(size_t(p.span_count));
    std::vector<std::vector<Stereo>> filtered(size_t(p.span_count));
    for (const auto& chain : chains(p)) {
        if (chain.empty())
            continue;
        const auto& head = p.spans[chain.front()];
        if (head.variable_pitch) {
            double position = head.position;
            int previousLoopEnd = 0, previousLoopStart = 0;
            for (int s : chain) {
                const auto& v = p.spans[s];
                if (!v.loop_end && previousLoopEnd > 0 && position >= previousLoopEnd)
                    position = previousLoopStart + std::fmod(position - previousLoopStart,
                                                             double(previousLoopEnd - previousLoopStart));
                previousLoopEnd = v.loop_end;
                previousLoopStart = v.loop_first;
                const double slope = std::exp(v.pitch_slope);
                double step = v.step;
                auto& out = phase[size_t(s)];
                out.resize(size_t(std::max(0, v.end - v.begin)));
                for (int f = v.begin; f < v.end; ++f) {
                    out[size_t(f - v.begin)] = position;
                    double cents = 0, depth = v.extra_envelope[1].depth;
                    for (int j = v.first_modulation; j < v.first_modulation + v.modulation_count; ++j) {
                        const auto& m = p.modulation[j];
                        if (m.target == CLAY_MOD_PITCH)
                            cents += job.modulationAt(m, f);
                        if (m.target == CLAY_MOD_PITCHEG_DEPTH)
                            depth += job.modulationAt(m, f);
                    }
                    if (v.active_envelope[1])
                        cents += depth * envelope(v.extra_envelope[1], f);
                    position += step * std::pow(2.0, std::min(24000.0, std::max(-24000.0, cents)) / 1200);
                    step *= slope;
                }
            }
        }
        // Filter source: voice audio before pan (mono) times the job gain (render.cu:195-213).
        for (int s : chain) {
            const auto& v = p.spans[s];
            auto& out = filtered[size_t(s)];
            out.resize(size_t(std::max(0, v.end - v.begin)));
            for (int f = v.begin; f < v.end; ++f) {
                const double ph = v.variable_pitch ? phase[size_t(s)][size_t(f - v.begin)] : -1;
                const Stereo a = R.voiceAudio(job, v, job.panBase(f), f, true, ph);
                const double g = job.jobGain(f);
                out[size_t(f - v.begin)] = {a.l * g, a.r * g};
            }
        }
        // Filter recurrence across the chain (render.cu:338-448).
        bool hasFilters = false;
        for (const auto& f : head.filters)
            hasFilters |= f.type != CLAY_FILTER_NONE;
        if (!hasFilters)
            continue;
        double z1[CLAY_FILTER_SLOTS][3][2] = {}, z2[CLAY_FILTER_SLOTS][3][2] = {};
        Coefficients current[CLAY_FILTER_SLOTS];
        bool initialized = false;
        for (int s : chain) {
            const auto& v = p.spans[s];
            const int rate = v.sample_rate;
            const double amount = 1 - std::exp(-1000.0 / rate);
            int count[CLAY_FILTER_SLOTS];
            for (int f = 0; f < CLAY_FILTER_SLOTS; ++f)
                count[f] = sections(v.filters[f].type);
            bool moving = v.active_envelope[0] != 0;
            for (int j = v.first_modulation; j < v.first_modulation + v.modulation_count; ++j) {
                const int t = p.modulation[j].target;
                moving |= (t >= CLAY_MOD_CUTOFF && t <= CLAY_MOD_RESONANCE2) || t == CLAY_MOD_FILEG_DEPTH ||
                          t >= CLAY_MOD_EQ1_FREQUENCY;
            }
            auto& buffer = filtered[size_t(s)];
            for (int fr = v.begin; fr < v.end; ++fr) {
                if (moving || !initialized) {
                    double cents[2] = {0, 0}, db[2] = {0, 0}, eq[3][3] = {};
                    double depth = v.extra_envelope[0].depth;
                    for (int j = v.first_modulation; j < v.first_modulation + v.modulation_count; ++j) {
                        const auto& m = p.modulation[j];
                        const int t = m.target;
                        if (t == CLAY_MOD_CUTOFF) cents[0] += job.modulationAt(m, fr);
                        if (t == CLAY_MOD_CUTOFF2) cents[1] += job.modulationAt(m, fr);
                        if (t == CLAY_MOD_RESONANCE) db[0] += job.modulationAt(m, fr);
                        if (t == CLAY_MOD_RESONANCE2) db[1] += job.modulationAt(m, fr);
                        if (t == CLAY_MOD_FILEG_DEPTH) depth += job.modulationAt(m, fr);
                        if (t >= CLAY_MOD_EQ1_FREQUENCY) {
                            const int id = t - CLAY_MOD_EQ1_FREQUENCY;
                            eq[id / 3][id % 3] += job.modulationAt(m, fr);
                        }
                    }
                    if (v.active_envelope[0])
                        cents[0] += depth * envelope(v.extra_envelope[0], fr);
                    for (int f = 0; f < CLAY_FILTER_SLOTS; ++f) {
                        if (!count[f])
                            continue;
                        const auto& flt = v.filters[f];
                        const Coefficients target = f < 2
                            ? coefficients(flt.type, flt.cutoff * std::pow(2.0, cents[f] / 1200), flt.resonance + db[f], rate, 0, 0)
                            : coefficients(flt.type, flt.cutoff + eq[f - 2][0], 0, rate, flt.bandwidth + eq[f - 2][1],
                                           flt.gain + eq[f - 2][2]);
                        if (initialized) {
                            auto& c = current[f];
                            c.b0 += amount * (target.b0 - c.b0);
                            c.b1 += amount * (target.b1 - c.b1);
                            c.b2 += amount * (target.b2 - c.b2);
                            c.a1 += amount * (target.a1 - c.a1);
                            c.a2 += amount * (target.a2 - c.a2);
                        } else {
                            current[f] = target;
                        }
                    }
                    initialized = true;
                }
                auto& x = buffer[size_t(fr - v.begin)];
                double l = x.l, r = v.channels == 2 ? x.r : 0;
                for (int f = 0; f < CLAY_FILTER_SLOTS; ++f)
                    for (int sec = 0; sec < count[f]; ++sec) {
                        l = filterSample(l, current[f], z1[f][sec][0], z2[f][sec][0]);
                        if (v.channels == 2)
                            r = filterSample(r, current[f], z1[f][sec][1], z2[f][sec][1]);
                    }
                x.l = l;
                x.r = v.channels == 2 ? r : l;
            }
        }
    }
    // Voice-silence policy: retirement frame per chain.
    std::vector<int> silence;
    for (const auto& chain : chains(p))
        silence.push_back(chainSilence(p, chain, filtered, R.silence));
    if (tails)
        *tails = silence;
    // Output mix (render.cu:458-492): dry voices times the job gain plus the wet chains.
    for (int s = 0; s < p.span_count; ++s) {
        const auto& v = p.spans[s];
        for (int f = v.begin; f < v.end; ++f) {
            if (v.filter_track >= 0) {
                if (f >= silence[size_t(v.filter_track)])
                    continue;
                Stereo x = filtered[size_t(s)][size_t(f - v.begin)];
                if (v.channels == 1) {
                    double offset = job.panBase(f);
                    for (int j = v.first_modulation; j < v.first_modulation + v.modulation_count; ++j)
                        if (p.modulation[j].target == CLAY_MOD_PAN)
                            offset += job.modulationAt(p.modulation[j], f);
                    const Stereo g = Renderer::panGains(v, offset);
                    x = {x.l * g.l, x.l * g.r};
                }
                wet[size_t(f)].l += x.l;
                wet[size_t(f)].r += x.r;
            } else {
                const Stereo a = R.voiceAudio(job, v, job.panBase(f), f, false, -1);
                dry[size_t(f)].l += a.l;
                dry[size_t(f)].r += a.r;
            }
        }
    }
    mix.resize(size_t(frames));
    for (int f = 0; f < frames; ++f) {
        const double g = job.jobGain(f);
        mix[size_t(f)] = {dry[size_t(f)].l * g + wet[size_t(f)].l, dry[size_t(f)].r * g + wet[size_t(f)].r};
    }
}

int16_t toPcm(double x)  // clamp then round half even (render.cu:102-105, tail.cu:162-163)
{
    return int16_t(std::nearbyint(std::min(32767.0, std::max(-32768.0, x))));
}

// Ties-to-even average of two PCM16 values (src/cuda/tail.cu:164-165).
int monoOf(int left, int right)
{
    const int sum = left + right;
    return (sum >> 1) + ((sum & 1) & ((sum >> 1) & 1));
}

// Half-cosine fade of one sample (src/cuda/tail.cu:46-51).
int16_t fadeSample(int value, int index, int count)
{
    const double gain = count == 1 ? 0 : .5 * (1 + std::cos(kPi * index / (count - 1)));
    return int16_t(std::nearbyint(value * gain));
}

struct Finished {
    clay_tail_result result{};
    std::vector<int16_t> pcm;  // dataFrames x channels
};

struct TailConfig {
    int window, quietFrames, padding, fade;
    double relative, floor, ceiling;
};

// Natural tail (src/cuda/tail.cu:54-213) on one job's mix.
void naturalTail(const clay_plan_view& p, const std::vector<Stereo>& mix, const TailConfig& tail, Finished& out)
{
    auto& res = out.result;
    res = {};
    const int frames = p.frames;
    double peak = 0;
    int end = 0;
    for (int first = 0; first < frames; first += kTailBlock) {
        const int blockEnd = std::min(first + kTailBlock, frames);
        double power = 0, blockPeak = 0;
        bool finite = true;
        for (int f = first; f < blockEnd; ++f) {
            const auto& x = mix[size_t(f)];
            if (!std::isfinite(x.l) || !std::isfinite(x.r)) {
                finite = false;
                continue;
            }
            blockPeak = std::max(blockPeak, std::max(std::abs(x.l), std::abs(x.r)));
            power += x.l * x.l + x.r * x.r;
        }
        if (!finite) {
            res.status = CLAY_TAIL_NONFINITE;
            return;
        }
        peak = std::max(peak, blockPeak);
        const int limitEnd = std::min(p.limit, first + kTailBlock);
        if (limitEnd < p.music_end)
            continue;
        const double threshold = 1e-12 * (limitEnd - first) * 2;
        if (power <= threshold) {
            end = limitEnd;
            break;
        }
    }
    if (!end) {
        const int silent = (frames + kTailBlock - 1) / kTailBlock * kTailBlock;
        if (silent < p.limit)
            end = std::min(p.limit, std::max(silent + kTailBlock, (p.music_end + kTailBlock - 1) / kTailBlock * kTailBlock));
    }
    if (!end) {
        res.status = CLAY_TAIL_DURATION;
        return;
    }
    const double attenuation = peak >= 1 ? 20 * std::log10(peak) + 1 : 0;
    if (attenuation > 5 + 1e-6) {
        res.status = CLAY_TAIL_PEAK;
        return;
    }
    res.frames = end;
    res.data_frames = std::min(end, frames);
    res.scale = float(32768 * std::pow(10., -attenuation / 20));
    const double scale = res.scale;
    // Mono PCM16 with 50 ms window statistics (tail.cu:147-174).
    std::vector<int16_t> mono(size_t(res.data_frames));
    const int windows = (res.data_frames + tail.window - 1) / tail.window;
    std::vector<double> wPower(size_t(windows), 0);
    std::vector<int> wPeak(size_t(windows), 0);
    for (int f = 0; f < res.data_frames; ++f) {
        const int l = toPcm(mix[size_t(f)].l * scale), r = toPcm(mix[size_t(f)].r * scale);
        const int v = monoOf(l, r);
        mono[size_t(f)] = int16_t(v);
        wPower[size_t(f / tail.window)] += double(v) * v;
        wPeak[size_t(f / tail.window)] = std::max(wPeak[size_t(f / tail.window)], std::abs(v));
    }
    // Weak-tail trim (tail.cu:183-207).
    double maxPower = 0;
    for (double w : wPower)
        maxPower = std::max(maxPower, w);
    const double threshold = std::min(tail.ceiling, std::max(tail.floor, std::sqrt(maxPower / tail.window) * tail.relative));
    for (int i = windows - 1; i >= 0; --i) {
        if (std::sqrt(wPower[size_t(i)] / tail.window) > threshold || wPeak[size_t(i)] > threshold * 4) {
            const int lastSound = std::min(res.frames, (i + 1) * tail.window);
            if (res.frames - lastSound >= tail.quietFrames) {
                res.frames = lastSound + tail.padding;
                res.data_frames = std::min(res.frames, res.data_frames);
                res.fade = std::min(res.frames, tail.fade);
            }
            break;
        }
    }
    out.pcm.assign(mono.begin(), mono.begin() + res.data_frames);
}

// Fade, peak and audible extent over the packed PCM (tail.cu:216-297).
void finalizePcm(Finished& out, int channels, bool countClipped)
{
    auto& res = out.result;
    res.peak = 0;
    res.first = res.frames;
    res.last = -1;
    const int fadeStart = res.frames - res.fade;
    for (int f = 0; f < res.data_frames; ++f) {
        int framePeak = 0;
        for (int c = 0; c < channels; ++c) {
            auto& v = out.pcm[size_t(f) * size_t(channels) + size_t(c)];
            if (f >= fadeStart)
                v = fadeSample(v, f - fadeStart, res.fade);
            framePeak = std::max(framePeak, std::abs(int(v)));
            if (countClipped)
                res.audio.clipped += std::abs(int(v)) >= 32767;
        }
        res.peak = std::max(res.peak, framePeak);
        if (framePeak > 3) {
            res.first = std::min(res.first, f);
            res.last = f;
        }
    }
}

using Power = unsigned long long;

// Body windows: 250 ms windows inside 2 s chunks (quality.cu:13-19).
int bodyBlockStart(int block, const clay_audio_config& a)
{
    const int chunk = 2 * a.rate, span = 5 * a.window;
    const int windows = (chunk + span - 1) / span;
    return (block / windows) * chunk + (block % windows) * span;
}

Power monoPower(const std::vector<int16_t>& pcm, int channels, int frame)
{
    const long long v = channels == 1 ? 2LL * pcm[size_t(frame)]
                                      : (long long)pcm[2 * size_t(frame)] + pcm[2 * size_t(frame) + 1];
    return Power(v * v);
}

double level(Power power, int window)
{
    return power ? std::max(-120.0, 10 * std::log10(double(power) / (window * 4294967296.))) : -120.0;
}

// Quality measures and verdict (quality.cu:21-164).
void assess(const clay_plan_view& p, const clay_audio_config& a, int channels, Finished& out)
{
    auto& res = out.result;
    if (res.status != CLAY_TAIL_OK)
        return;
    auto& st = res.audio;
    const int body = a.enabled || p.volume_group ? p.body_frames : 0;
    const int end = std::min(res.frames, body);
    Power total = 0;
    int silent = 0, run = 0;
    for (int w = 0, first = 0; first < end; ++w) {
        const int next = std::min(end, bodyBlockStart(w + 1, a));
        const int stop = std::min(std::min(body, res.data_frames), bodyBlockStart(w + 1, a));
        Power wp = 0;
        for (int f = bodyBlockStart(w, a); f < stop; ++f)
            wp += monoPower(out.pcm, channels, f);
        total += wp;
        if (a.enabled) {
            if (double(wp) < (next - first) * a.silence_power) {
                silent += next - first;
                run += next - first;
                st.longest_silence = std::max(run, st.longest_silence);
            } else {
                run = 0;
            }
        }
        first = next;
    }
    st.body_rms = body ? std::sqrt(double(total) / body) * .5 : 0;
    if (!a.enabled)
        return;
    st.measured = 1;
    st.median = st.quartile = -120;
    st.silent_fraction = double(silent) / std::max(1, end);
    st.clipped_fraction = double(st.clipped) / (double(std::max(1, res.frames)) * channels);
    std::vector<Power> onset;
    for (int i = 0; i < p.onset_count; ++i) {
        Power best = 0;
        for (int w = 0; w < 5; ++w) {
            const int first = p.onsets[i] + w * a.window;
            const int stop = std::min(first + a.window, res.data_frames);
            Power wp = 0;
            for (int f = first; f < stop; ++f)
                wp += monoPower(out.pcm, channels, f);
            best = std::max(best, wp);
        }
        onset.push_back(best);
    }
    if (!onset.empty()) {
        std::sort(onset.begin(), onset.end());
        const auto percentile = [&](double fraction) {
            const double position = (onset.size() - 1) * fraction;
            const int first = int(position);
            const double lower = level(onset[size_t(first)], a.window);
            return size_t(first + 1) < onset.size()
                ? lower + (level(onset[size_t(first) + 1], a.window) - lower) * (position - first) : lower;
        };
        st.median = percentile(.5);
        st.quartile = percentile(.25);
    }
    if (res.frames > a.max_frames) res.status = CLAY_TAIL_AUDIO_DURATION;
    else if (onset.empty() || st.median < a.min_median || st.quartile < a.min_quartile) res.status = CLAY_TAIL_QUIET_PERFORMANCE;
    else if (st.silent_fraction > a.max_silent_fraction || st.longest_silence > a.max_silent_frames) res.status = CLAY_TAIL_EXCESSIVE_SILENCE;
    else if (res.peak < a.min_peak) res.status = CLAY_TAIL_SILENT_AUDIO;
    else if (st.clipped_fraction > a.max_clipped_fraction) res.status = CLAY_TAIL_CLIPPING;
    st.verdict = res.status;
}

// Shared group gain (quality.cu:172-209) over consecutive jobs of one group.
void matchGroup(std::vector<Finished>& jobs, size_t first, size_t end, int minimum)
{
    double desired = 0, safe = 1e30;
    int accepted = 0;
    for (size_t i = first; i < end; ++i) {
        auto& r = jobs[i].result;
        if (r.status != CLAY_TAIL_OK)
            continue;
        const double rms = r.audio.body_rms;
        if (rms <= 0 || r.last < 0) {
            r.status = CLAY_TAIL_SILENT_BODY;
            continue;
        }
        desired += rms;
        safe = std::min(safe, 32767 * .8912509381337456 * rms / std::max(1, r.peak));
        ++accepted;
    }
    const double target = std::min(desired / std::max(1, accepted), safe);
    for (size_t i = first; i < end; ++i) {
        auto& r = jobs[i].result;
        if (r.status != CLAY_TAIL_OK)
            continue;
        if (accepted < minimum) {
            r.status = CLAY_TAIL_INCOMPLETE_GROUP;
        } else {
            r.audio.target_rms = target;
            r.audio.gain = target / r.audio.body_rms;
        }
    }
}

// Gain and compaction of accepted jobs (quality.cu:211-274).
void packAccepted(Finished& job, int channels)
{
    auto& r = job.result;
    if (r.status != CLAY_TAIL_OK || r.last < 0)
        r.data_frames = 0;
    job.pcm.resize(size_t(r.data_frames) * size_t(channels));
    const double gain = r.audio.gain;
    if (!(gain > 0))
        return;
    r.peak = 0;
    r.first = r.frames;
    r.last = -1;
    for (int f = 0; f < r.data_frames; ++f) {
        int framePeak = 0;
        for (int c = 0; c < channels; ++c) {
            auto& v = job.pcm[size_t(f) * size_t(channels) + size_t(c)];
            v = int16_t(std::nearbyint(v * gain));
            framePeak = std::max(framePeak, std::abs(int(v)));
        }
        r.peak = std::max(r.peak, framePeak);
        if (framePeak > 3) {
            r.first = std::min(r.first, f);
            r.last = f;
        }
    }
}

} // namespace

extern "C" {

const char* clay_oracle_last_error(void) { return g_error.c_str(); }

int clay_oracle_sinc_table(int taps, float* out, size_t capacity)
{
    if (taps < 2 || taps > CLAY_MAX_SINC_TAPS || taps % 2 || capacity < size_t(taps) * (kPhases + 1))
        return CLAY_E_INVALID;
    const auto t = sincTable(taps);
    std::memcpy(out, t.data(), t.size() * sizeof(float));
    return CLAY_OK;
}

float silenceLevel(float db) { return db < 0 ? float(std::pow(10.0, double(db) / 20.0)) : 0.f; }

int clay_oracle_voice_tails(const clay_oracle_config* cfg, const int16_t* bank, uint64_t bank_count,
                            const clay_plan_view* plan, int32_t* silence, size_t capacity)
{
    if (!cfg || !plan || !silence || capacity < size_t(std::max(0, plan->filter_tracks)))
        return CLAY_E_INVALID;
    Renderer R{bank, bank_count, cfg->sinc_taps, sincTable(cfg->sinc_taps)};
    R.silence = silenceLevel(cfg->voice_silence_db);
    Job job(plan);
    std::vector<Stereo> m;
    std::vector<int> tails;
    mixJob(R, job, m, &tails);
    std::copy(tails.begin(), tails.end(), silence);
    return CLAY_OK;
}

int clay_oracle_mix(const clay_oracle_config* cfg, const int16_t* bank, uint64_t bank_count, const clay_plan_view* plan,
                    double* mix, uint64_t capacity_frames)
{
    if (!cfg || !plan || (uint64_t)std::max(0, plan->frames) > capacity_frames)
        return CLAY_E_INVALID;
    Renderer R{bank, bank_count, cfg->sinc_taps, sincTable(cfg->sinc_taps)};
    R.silence = silenceLevel(cfg->voice_silence_db);
    Job job(plan);
    std::vector<Stereo> m;
    mixJob(R, job, m);
    for (size_t f = 0; f < m.size(); ++f) {
        mix[2 * f] = m[f].l;
        mix[2 * f + 1] = m[f].r;
    }
    return CLAY_OK;
}

int clay_oracle_render(const clay_oracle_config* cfg, const int16_t* bank, uint64_t bank_count,
                       const clay_plan_view* plans, size_t count, clay_tail_result* results, int16_t* pcm,
                       uint64_t pcm_capacity, uint64_t* pcm_samples)
{
    try {
        const bool natural = cfg->tail_policy == CLAY_POLICY_NATURAL;
        const int channels = natural ? 1 : 2;
        const auto& a = cfg->audio;
        const int rate = a.rate;
        const TailConfig tail{a.window, 2 * rate, int(std::nearbyint(rate / 4.)), int(std::nearbyint(rate / 10.)),
                              std::pow(10., -1.5), 32768 * std::pow(10., -3.9), 32768 * std::pow(10., -3.)};
        Renderer R{bank, bank_count, cfg->sinc_taps, sincTable(cfg->sinc_taps)};
        R.silence = silenceLevel(cfg->voice_silence_db);
        std::vector<Finished> jobs(count);
        std::atomic<size_t> next{0};
        const auto work = [&] {
            for (size_t i; (i = next++) < count;) {
                const auto& p = plans[i];
                Job job(&p);
                std::vector<Stereo> mix;
                mixJob(R, job, mix);
                auto& out = jobs[i];
                if (natural) {
                    naturalTail(p, mix, tail, out);
                } else {
                    out.result = {};
                    out.result.frames = out.result.data_frames = p.frames;
                    out.result.fade = p.fade_frames;
                    out.pcm.resize(size_t(p.frames) * 2);
                    for (int f = 0; f < p.frames; ++f) {
                        out.pcm[2 * size_t(f)] = toPcm(mix[size_t(f)].l * 32768);
                        out.pcm[2 * size_t(f) + 1] = toPcm(mix[size_t(f)].r * 32768);
                    }
                }
                if (out.result.status != CLAY_TAIL_OK) {
                    out.pcm.clear();
                    out.result.data_frames = 0;
                    out.result.first = out.result.frames;
                    out.result.last = -1;
                    continue;
                }
                finalizePcm(out, channels, a.enabled != 0);
                if (a.enabled || cfg->group_minimum)
                    assess(p, a, channels, out);
            }
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < std::max(1, cfg->threads); ++t)
            pool.emplace_back(work);
        work();
        for (auto& t : pool)
            t.join();
        const bool assessing = a.enabled || cfg->group_minimum;
        if (cfg->group_minimum)
            for (size_t first = 0, end; first < count; first = end) {
                end = first + 1;
                if (!plans[first].volume_group)
                    continue;
                while (end < count && plans[end].volume_group == plans[first].volume_group)
                    ++end;
                matchGroup(jobs, first, end, cfg->group_minimum);
            }
        uint64_t offset = 0;
        for (size_t i = 0; i < count; ++i) {
            auto& j = jobs[i];
            if (assessing)
                packAccepted(j, channels);
            j.result.offset = offset;
            offset += uint64_t(j.result.data_frames);
        }
        const uint64_t samples = offset * uint64_t(channels);
        if (pcm_samples)
            *pcm_samples = samples;
        if (pcm && samples > pcm_capacity) {
            g_error = "pcm capacity too small";
            return CLAY_E_INVALID;
        }
        for (size_t i = 0; i < count; ++i) {
            results[i] = jobs[i].result;
            if (pcm && !jobs[i].pcm.empty())
                std::memcpy(pcm + jobs[i].result.offset * uint64_t(channels), jobs[i].pcm.data(),
                            size_t(jobs[i].result.data_frames) * size_t(channels) * sizeof(int16_t));
        }
        return CLAY_OK;
    } catch (const std::exception& e) {
        g_error = e.what();
        return CLAY_E_INVALID;
    }
}

double clay_oracle_envelope(const clay_voice_span* span, int frame) { return envelope(*span, frame); }

int clay_oracle_filter_coefficients(int type, double cutoff, double resonance, int rate, double bandwidth, double gain,
                                    double out[5])
{
    const auto c = coefficients(type, cutoff, resonance, rate, bandwidth, gain);
    out[0] = c.b0, out[1] = c.b1, out[2] = c.b2, out[3] = c.a1, out[4] = c.a2;
    return CLAY_OK;
}

int clay_oracle_interpolate(const clay_voice_span* span, const int16_t* bank, int taps, double position, double out[2])
{
    Renderer R{bank, 0, taps, sincTable(taps)};
    const auto s = R.interpolateAt(*span, position);
    out[0] = s.l;
    out[1] = s.r;
    return CLAY_OK;
}

int clay_oracle_mono(int left, int right) { return monoOf(left, right); }

} // extern "C"
