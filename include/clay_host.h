/*
 * clay_host.h - C ABI of the host side of the render core: SFZ and MIDI
 * loading and the CPU planner that turns one job (SFZ + MIDI) into the
 * clay_plan_view the HIP core renders.
 *
 * Replaces the planning window of the reference CLI (src/main.cpp:298-361:
 * loadInstrument, loadMidi, planRender, musicEnd/limit/fade) and the sample
 * bank it fills (src/io/audio.h:18-37). The bank is int16.
 */
#ifndef CLAY_HOST_H
#define CLAY_HOST_H

#include "clay_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Planner options shared by the jobs of one session (src/main.cpp:48-69). */
typedef struct clay_session_config {
    int32_t rate;                 /* --samplerate (default 44100) */
    int32_t tail_policy;          /* enum clay_tail_policy */
    int32_t measure;              /* QA onsets/body frames (--quality-gate or --group-volume) */
    int32_t keep_onsets;          /* copy onsets into plans (--quality-gate != off) */
    double tail_max_seconds;      /* custom: --tail-max-seconds (< 0 = unset) */
    double max_audio_seconds;     /* natural: --max-audio-seconds (< 0 = unset) */
    int32_t tail_fade_ms;         /* custom: --tail-fade-ms (< 0 = default 100) */
} clay_session_config;

typedef struct clay_session clay_session;

int clay_session_create(const clay_session_config* config, clay_session** out);
int clay_session_destroy(clay_session* s);

/* Plan one job and append it to the session. On success *index receives the
 * job index; on a job-level failure (bad SFZ/MIDI, no playable notes) the
 * call returns CLAY_E_INVALID and clay_session_last_error holds the message
 * the reference prints on its REJECT line. `group` = volume group id (0 = none). */
int clay_session_add_job(clay_session* s, const char* sfz_path, const char* midi_path, uint64_t group, size_t* index);

/* Plan view of job `index`; valid until the session is destroyed or reset. */
const clay_plan_view* clay_session_plan(clay_session* s, size_t index);
size_t clay_session_job_count(const clay_session* s);

/* Decoded int16 sample bank (grows as jobs decode samples). */
const int16_t* clay_session_bank(const clay_session* s, uint64_t* count);

/* Drop all jobs and the decoded bank (instruments stay described). */
int clay_session_reset(clay_session* s);

const char* clay_session_last_error(const clay_session* s);

/* Parsed MIDI events of one file (src/io/midi.cpp:331-342): events holds
 * (frame, kind, a, b) quadruples, kind 0 on / 1 off / 2 control / 3 bend.
 * Returns CLAY_E_INVALID with the loader's message in `error` on failure. */
int clay_midi_load(const char* path, int rate, int tail, int measure, int* events, int capacity, int* count, int* end,
                   int* body_frames, int* onsets, int onset_capacity, int* onset_count, char* error, int error_capacity);

/* sizeof() of an ABI struct by its C name (0 if unknown), so bindings in
 * other languages can check their struct layouts. */
size_t clay_abi_sizeof(const char* name);

/* PCM16 WAV/FLAC writers (src/io/audio.cpp:264-285, :290-316). */
int clay_write_wave(const char* path, const int16_t* pcm, int32_t data_frames, int32_t frames, int32_t rate,
                    int32_t channels);
int clay_write_flac(const char* path, const int16_t* pcm, int32_t data_frames, int32_t frames, int32_t rate,
                    int32_t channels, int32_t level);
/* Decode a FLAC file into `pcm` (capacity in samples); *samples = total samples. */
int clay_read_flac(const char* path, int32_t* rate, int32_t* channels, int16_t* pcm, uint64_t capacity, uint64_t* samples);

#ifdef __cplusplus
}
#endif
#endif
