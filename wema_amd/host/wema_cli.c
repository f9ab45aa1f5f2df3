/*
 * wema - command-line front end of the MI355X pipeline with the option set of the
 * reference's CLI (src/main.c:16-218): the same options, defaults, checks and messages'
 * meaning, plus --levels (the DWT level count, which the reference fixes by its automatic
 * rule, src/dwt.c:588-596) and --device (the HIP device). FFmpeg does the decoding and
 * encoding in child processes, as in the reference (src/ffmpeg_io.c:49-262): ffprobe for
 * the stream's size, frame rate and length; one ffmpeg writing RGB24 frames to a pipe; one
 * ffmpeg reading the output frames from a pipe (with the input's audio copied across).
 *
 * Frames stay RGB24 from pipe to pipe: they are read into pinned host buffers, queued to the
 * device session with their frame number (wema_dev_submit: uploads, the pipeline and
 * downloads run on their own streams, two batches on the device at a time) and collected in
 * order for the encoder. The reference converts each byte to a float and back around its
 * pipeline (src/ffmpeg_io.c:292-294, 304-310); the device does the same conversions, so the
 * output bytes are those of the float path. As in the reference (src/main.c:355), the luma
 * IIR mode runs with the batch path's semantics (the guided filter always on) and the other
 * modes with the per-frame path's.
 */
#define _POSIX_C_SOURCE 200809L
#include <errno.h>
#include <signal.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/types.h>
#include <sys/wait.h>
#include <unistd.h>

#include "wema_dev.h"

typedef struct {
    const char *input, *output, *codec, *options;
    float amp, f_low, f_high;
    int window, verbose, edge_aware, bilateral, color, use_iir, batch, threads, levels, device;
} Options;

static void usage(const char *prog) {
    fprintf(stderr,
            "WEMA - Wavelet-based Eulerian Motion Amplification (MI355X)\n"
            "\n"
            "Usage: %s -i <input> [-o <output>] [options]\n"
            "\n"
            "Options:\n"
            "  -i, --input <file>     Input video file (required)\n"
            "  -o, --output <file>    Output video file (default: input.wema.mkv)\n"
            "  -a, --amp <factor>     Amplification factor (default: 120)\n"
            "  --fl <freq>            Low frequency cutoff in Hz (default: 0.5)\n"
            "  --fh <freq>            High frequency cutoff in Hz (default/max: half of framerate)\n"
            "  --ff-codec <codec>     FFmpeg video codec (default: ffv1)\n"
            "  --ff-option <opts>     FFmpeg encoder options\n"
            "  --temporal-window <n>  Temporal window size (default: 32, min:4, max: 256)\n"
            "  --no-edge-aware        Disable edge-aware guided filter\n"
            "  --bilateral-filter     Enable bilateral temporal filtering\n"
            "  --color                Enable color mode (amplify color channels)\n"
            "  --haar                 Sliding temporal window instead of the IIR band-pass\n"
            "  --batch <n>            Frames per device batch (default: 64)\n"
            "  --threads <n>          Accepted for compatibility (the GPU does the work)\n"
            "  --levels <n>           DWT levels, 1-12 (default: automatic, at most 4)\n"
            "  --device <n>           HIP device (default: 0)\n"
            "  -v, --verbose          Verbose output\n"
            "  -h, --help             Show this help\n",
            prog);
}

/* The next argument of an option, or NULL with the reference's message. */
static const char *value_of(int argc, char **argv, int *i) {
    if (++*i >= argc) {
        fprintf(stderr, "Error: %s requires an argument\n", argv[*i - 1]);
        return NULL;
    }
    return argv[*i];
}

static int parse(int argc, char **argv, Options *o) {
    memset(o, 0, sizeof(*o));
    o->amp = 120.0f;          /* defaults of src/main.c:65-75 */
    o->f_low = 0.5f;
    o->f_high = 1048576.0f;
    o->window = 32;
    o->edge_aware = 1;
    o->use_iir = 1;
    o->batch = 64;
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i], *v;
#define IS(s) (strcmp(a, s) == 0)
        if (IS("-i") || IS("--input")) {
            if (!(v = value_of(argc, argv, &i))) return -1;
            o->input = v;
        } else if (IS("-o") || IS("--output")) {
            if (!(v = value_of(argc, argv, &i))) return -1;
            o->output = v;
        } else if (IS("-a") || IS("--amp")) {
            if (!(v = value_of(argc, argv, &i))) return -1;
            o->amp = strtof(v, NULL);
            if (!(o->amp > 0)) { fprintf(stderr, "Error: amplification factor must be positive\n"); return -1; }
        } else if (IS("--fl")) {
            if (!(v = value_of(argc, argv, &i))) return -1;
            o->f_low = strtof(v, NULL);
            if (o->f_low < 0) { fprintf(stderr, "Error: low frequency must be non-negative\n"); return -1; }
        } else if (IS("--fh")) {
            if (!(v = value_of(argc, argv, &i))) return -1;
            o->f_high = strtof(v, NULL);
            if (!(o->f_high > 0)) { fprintf(stderr, "Error: high frequency must be positive\n"); return -1; }
        } else if (IS("--temporal-window")) {
            if (!(v = value_of(argc, argv, &i))) return -1;
            o->window = atoi(v);
            if (o->window < 4) { fprintf(stderr, "Error: temporal window must be at least 4\n"); return -1; }
            if (o->window > 256) { fprintf(stderr, "Error: temporal window cannot exceed 256\n"); return -1; }
        } else if (IS("--ff-codec")) {
            if (!(v = value_of(argc, argv, &i))) return -1;
            o->codec = v;
        } else if (IS("--ff-option")) {
            if (!(v = value_of(argc, argv, &i))) return -1;
            o->options = v;
        } else if (IS("-v") || IS("--verbose")) {
            o->verbose = 1;
        } else if (IS("--no-edge-aware")) {
            o->edge_aware = 0;
        } else if (IS("--bilateral-filter")) {
            o->bilateral = 1;
        } else if (IS("--color")) {
            o->color = 1;
        } else if (IS("--haar")) {
            o->use_iir = 0;
        } else if (IS("--batch")) {
            if (!(v = value_of(argc, argv, &i))) return -1;
            o->batch = atoi(v);
            if (o->batch < 1) { fprintf(stderr, "Error: batch size must be at least 1\n"); return -1; }
            if (o->batch > 256) { fprintf(stderr, "Error: batch size cannot exceed 256\n"); return -1; }
        } else if (IS("--threads")) {
            if (!(v = value_of(argc, argv, &i))) return -1;
            o->threads = atoi(v);
            if (o->threads < 0) { fprintf(stderr, "Error: thread count must be non-negative\n"); return -1; }
        } else if (IS("--levels")) {
            if (!(v = value_of(argc, argv, &i))) return -1;
            o->levels = atoi(v);
            if (o->levels < 1 || o->levels > 12) { fprintf(stderr, "Error: levels must be in 1..12\n"); return -1; }
        } else if (IS("--device")) {
            if (!(v = value_of(argc, argv, &i))) return -1;
            o->device = atoi(v);
            if (o->device < 0) { fprintf(stderr, "Error: device must be non-negative\n"); return -1; }
        } else if (IS("-h") || IS("--help")) {
            usage(argv[0]);
            exit(0);
        } else {
            fprintf(stderr, "Error: Unknown option: %s\n", a);
            return -1;
        }
#undef IS
    }
    if (!o->input) { fprintf(stderr, "Error: Input file required (-i)\n"); return -1; }
    if (o->f_low >= o->f_high) { fprintf(stderr, "Error: Low frequency must be less than high frequency\n"); return -1; }
    return 0;
}

/* input.ext -> input.wema.mkv (src/main.c:47-61) */
static char *default_output(const char *in) {
    const size_t n = strlen(in);
    const char *dot = strrchr(in, '.');
    const size_t base = dot ? (size_t)(dot - in) : n;
    char *out = malloc(base + 16);
    if (!out) return NULL;
    memcpy(out, in, base);
    strcpy(out + base, ".wema.mkv");
    return out;
}

/* ffprobe's width, height, frame rate and frame count of the first video stream. The path
 * is passed as an argument (no shell). */
static int probe(const char *path, int *w, int *h, float *fps, long *frames) {
    int fd[2];
    if (pipe(fd) < 0) return -1;
    const pid_t pid = fork();
    if (pid < 0) return -1;
    if (pid == 0) {
        dup2(fd[1], 1);
        close(fd[0]);
        close(fd[1]);
        execlp("ffprobe", "ffprobe", "-v", "quiet", "-select_streams", "v:0", "-show_entries",
               "stream=width,height,r_frame_rate,nb_frames", "-of", "csv=p=0", path, (char *)NULL);
        _exit(127);
    }
    close(fd[1]);
    char buf[512];
    size_t got = 0;
    ssize_t r;
    while (got + 1 < sizeof(buf) && (r = read(fd[0], buf + got, sizeof(buf) - 1 - got)) > 0) got += (size_t)r;
    buf[got] = 0;
    close(fd[0]);
    int status = 0;
    waitpid(pid, &status, 0);
    int num = 0, den = 1;
    char rate[64] = "", count[64] = "";
    if (sscanf(buf, "%d,%d,%63[^,\n],%63[^,\n]", w, h, rate, count) < 3) return -1;
    if (strchr(rate, '/')) sscanf(rate, "%d/%d", &num, &den);
    else num = atoi(rate);
    *frames = (count[0] && strcmp(count, "N/A") != 0) ? atol(count) : -1;
    if (*w <= 0 || *h <= 0 || num <= 0 || den <= 0) return -1;
    *fps = (float)num / (float)den;
    return 0;
}

/* ffmpeg as a child with one end of a pipe as its stdout (decoder) or stdin (encoder). */
static pid_t spawn(char *const *args, int *fd_out, int child_stdout) {
    int fd[2];
    if (pipe(fd) < 0) return -1;
    const pid_t pid = fork();
    if (pid < 0) return -1;
    if (pid == 0) {
        dup2(child_stdout ? fd[1] : fd[0], child_stdout ? 1 : 0);
        close(fd[0]);
        close(fd[1]);
        execvp(args[0], args);
        _exit(127);
    }
    close(child_stdout ? fd[1] : fd[0]);
    *fd_out = child_stdout ? fd[0] : fd[1];
    return pid;
}

static int read_full(int fd, uint8_t *p, size_t n) {
    size_t got = 0;
    while (got < n) {
        const ssize_t r = read(fd, p + got, n - got);
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) return got == 0 ? 0 : -1;   /* 0: end of stream; -1: a partial frame */
        got += (size_t)r;
    }
    return 1;
}

static int write_full(int fd, const uint8_t *p, size_t n) {
    while (n) {
        const ssize_t r = write(fd, p, n);
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) return -1;
        p += r;
        n -= (size_t)r;
    }
    return 0;
}

int main(int argc, char **argv) {
    Options o;
    if (parse(argc, argv, &o) < 0) {
        usage(argv[0]);
        return 1;
    }
    char *auto_out = NULL;
    if (!o.output) {
        if (!(auto_out = default_output(o.input))) return 1;
        o.output = auto_out;
    }
    signal(SIGPIPE, SIG_IGN);
    int W = 0, H = 0;
    float fps = 0.0f;
    long total = -1;
    if (probe(o.input, &W, &H, &fps, &total) < 0) {
        fprintf(stderr, "Error: Cannot open input: %s\n", o.input);
        free(auto_out);
        return 1;
    }
    if (o.f_high > fps / 2.0f - 0.0001f) o.f_high = fps / 2.0f - 0.0001f;   /* src/main.c:247-250 */
    const int batch_path = o.use_iir && !o.color;
    if (o.verbose) {
        fprintf(stderr, "Input: %s\n  Size: %dx%d\n  FPS: %.3f\n", o.input, W, H, fps);
        if (total > 0) fprintf(stderr, "  Frames: %ld\n", total);
        fprintf(stderr, "Output: %s\nSettings:\n  Amplification: %.1f\n  Frequency band: %.2f - %.2f Hz\n"
                        "  Temporal window: %d frames\n  Temporal filter: %s\n  Bilateral temporal: %s\n"
                        "  Edge-aware filter: %s\n  Color mode: %s\n  Levels: %d%s\n  Batch size: %d frames\n"
                        "  Device: %d\n",
                o.output, o.amp, o.f_low, o.f_high, o.window, o.use_iir ? "IIR (fast)" : "sliding window",
                o.bilateral ? "enabled" : "disabled", o.edge_aware || batch_path ? "enabled" : "disabled",
                o.color ? "enabled" : "disabled", o.levels ? o.levels : wema_dev_auto_levels(W, H),
                o.levels ? "" : " (automatic)", o.batch, o.device);
    }

    wema_dev_params p;
    wema_dev_params_default(&p, W, H, fps);
    p.amp = o.amp;
    p.f_low = o.f_low;
    p.f_high = o.f_high;
    p.window = o.window;
    p.levels = o.levels;
    p.use_iir = o.use_iir;
    p.color = o.color;
    p.edge_aware = o.edge_aware;
    p.bilateral = o.bilateral;
    p.batch_path = batch_path;
    p.max_batch = o.batch;
    p.device = o.device;
    /* decoder and encoder (src/ffmpeg_io.c:134-160, 190-245) */
    char size_s[32], fps_s[32];
    snprintf(size_s, sizeof(size_s), "%dx%d", W, H);
    snprintf(fps_s, sizeof(fps_s), "%.6f", fps);
    char *dec_args[] = {"ffmpeg", "-hide_banner", "-loglevel", "error", "-i", (char *)o.input, "-f", "rawvideo",
                        "-pix_fmt", "rgb24", "-", NULL};
    char *enc_args[64];
    int na = 0;
    enc_args[na++] = "ffmpeg";
    enc_args[na++] = "-hide_banner";
    enc_args[na++] = "-loglevel";
    enc_args[na++] = "error";
    enc_args[na++] = "-f";
    enc_args[na++] = "rawvideo";
    enc_args[na++] = "-pix_fmt";
    enc_args[na++] = "rgb24";
    enc_args[na++] = "-s";
    enc_args[na++] = size_s;
    enc_args[na++] = "-r";
    enc_args[na++] = fps_s;
    enc_args[na++] = "-i";
    enc_args[na++] = "-";
    enc_args[na++] = "-i";
    enc_args[na++] = (char *)o.input;
    enc_args[na++] = "-map";
    enc_args[na++] = "0:v";
    enc_args[na++] = "-map";
    enc_args[na++] = "1:a?";
    enc_args[na++] = "-c:a";
    enc_args[na++] = "copy";
    enc_args[na++] = "-c:v";
    enc_args[na++] = (char *)(o.codec ? o.codec : "ffv1");
    char *opts = o.options ? strdup(o.options) : NULL;   /* split on spaces, as a shell would */
    for (char *t = opts ? strtok(opts, " ") : NULL; t && na < 60; t = strtok(NULL, " ")) enc_args[na++] = t;
    enc_args[na++] = "-y";
    enc_args[na++] = (char *)o.output;
    enc_args[na] = NULL;
    int dec_fd = -1, enc_fd = -1;
    const pid_t dec = spawn(dec_args, &dec_fd, 1);
    const pid_t enc = dec > 0 ? spawn(enc_args, &enc_fd, 0) : -1;
    int rc = 0;
    if (dec <= 0 || enc <= 0) {
        fprintf(stderr, "Error: Cannot start ffmpeg\n");
        rc = 1;
    }
    /* the session is opened once the children run: nothing is started from this process
     * after it has touched the GPU */
    wema_dev_t *s = NULL;
    if (!rc && wema_dev_open(&s, &p) < 0) {
        fprintf(stderr, "Error: Failed to initialize WEMA: %s\n", wema_dev_last_error());
        rc = 1;
    }

    /* ring of pinned frame buffers: up to 2 batches queued on the device, the next being read */
    const size_t fb = (size_t)W * H * 3;
    const int ring = 3 * o.batch;
    uint8_t *in = rc ? NULL : wema_dev_host_alloc(fb * (size_t)ring);
    uint8_t *out = rc ? NULL : wema_dev_host_alloc(fb * (size_t)ring);
    if (!rc && (!in || !out)) {
        fprintf(stderr, "Error: %s\n", wema_dev_last_error());
        rc = 1;
    }
    long read_n = 0, written = 0;
    while (!rc) {
        const int r = read_full(dec_fd, in + fb * (size_t)(read_n % ring), fb);
        if (r < 0) { fprintf(stderr, "Error: truncated frame from the decoder\n"); rc = 1; break; }
        if (r == 0) break;
        if (wema_dev_submit(s, in + fb * (size_t)(read_n % ring), out + fb * (size_t)(read_n % ring), read_n) < 0) {
            fprintf(stderr, "Error: %s\n", wema_dev_last_error());
            rc = 1;
            break;
        }
        read_n++;
        while (read_n - written >= 2 * o.batch) {
            int64_t pts = -1;
            uint8_t *o8 = NULL;
            if (wema_dev_collect(s, &pts, &o8) < 0 || pts != written || write_full(enc_fd, o8, fb) < 0) {
                fprintf(stderr, "Error: %s\n", pts != written ? "frame order" : wema_dev_last_error());
                rc = 1;
                break;
            }
            written++;
        }
        if (o.verbose && read_n % 100 == 0) fprintf(stderr, "\rProcessed %ld frames", read_n);
    }
    while (!rc && written < read_n) {
        int64_t pts = -1;
        uint8_t *o8 = NULL;
        if (wema_dev_collect(s, &pts, &o8) < 0 || pts != written || write_full(enc_fd, o8, fb) < 0) {
            fprintf(stderr, "Error: writing frame %ld\n", written);
            rc = 1;
            break;
        }
        written++;
    }
    if (o.verbose) fprintf(stderr, "\rProcessed %ld frames\n", written);
    if (enc_fd >= 0) close(enc_fd);
    if (dec_fd >= 0) close(dec_fd);
    int st = 0;
    if (dec > 0) waitpid(dec, &st, 0);
    if (enc > 0) {
        waitpid(enc, &st, 0);
        if (!rc && !(WIFEXITED(st) && WEXITSTATUS(st) == 0)) {
            fprintf(stderr, "Error: the encoder failed\n");
            rc = 1;
        }
    }
    if (s) wema_dev_close(s);
    wema_dev_host_free(in);
    wema_dev_host_free(out);
    free(opts);
    free(auto_out);
    return rc;
}
