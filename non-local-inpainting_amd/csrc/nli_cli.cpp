// nli_cli.cpp -- the reference's command-line program (main.cpp) over libnli: the same
// usage, options and defaults, parameter checks and messages, method pairing, output file
// name, and the -showpyr / -shownnf dumps; every PatchMatch call and update runs on the
// GPU through the plugin classes of include/nli_plugins.hpp.
//
//   Inpainting input mask output [OPTIONS]
//
// Reference: main.cpp:43-61 (pick_option), :65-114 (options and usage), :115-147 (input,
// mask, depth), :160-217 (coarsest rate rule and parameter checks), :220-284 (parameters,
// pairing, process), :286-374 (output name, result, pyramid and NNF dumps).
//
// Differences, all deliberate:
// * errors print their message and exit with status 1 (the reference lets the
//   std::runtime_error escape main, which aborts);
// * the depth result is written as a grey image (the reference passes the one-channel
//   plane to write_rgb_image, which reads past its end, main.cpp:309);
// * PatchMatch draws from a counter-based generator seeded with time(NULL), as the
//   reference seeds rand() (main.cpp:273); -seed N fixes the seed instead;
// * -dryrun checks the options and inputs, prints the output file name and exits without
//   inpainting.
#include "nli_plugins.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>

using namespace nli;

namespace {

// Removes option -name (and, when default_value is given, the argument after it) from
// argv and returns its value; absent: default_value (for a flag without argument, the
// option's own text when present, null when absent).  main.cpp:43-61.
const char *pick_option(int *argc, char ***argv, const char *name, const char *default_value)
{
    const int takes = default_value ? 1 : 0;
    char **v = *argv;
    for (int i = 0; i < *argc - takes; i++) {
        if (v[i][0] != '-' || std::strcmp(v[i] + 1, name) != 0) continue;
        const char *value = takes ? v[i + 1] : v[i] + 1;
        const int n = *argc;
        for (int j = i; j < n - takes; j++) v[j] = v[j + takes + 1];   // argv[n] (NULL) moves down too
        *argc -= takes + 1;
        return value;
    }
    return default_value;
}

void usage(const char *prog, const std::string &method, int patch, int iters, int scales, float coarse, float conft,
           float confa, float lambda, const std::string &init, float psigma)
{
    std::fprintf(stderr, "USAGE:\n");
    std::fprintf(stderr, "%s input mask output [OPTIONS]\n\n", prog);
    std::fprintf(stderr, "Available options are:\n");
    std::fprintf(stderr, " -method \tmethod name [nlmeans/nlmedians/nlpoisson] (%s)\n", method.c_str());
    std::fprintf(stderr, " -patch  \tpatch side (%d)\n", patch);
    std::fprintf(stderr, " -iters  \tinpainting iterations (%d)\n", iters);
    std::fprintf(stderr, " -scales \tscales amount (%d)\n", scales);
    std::fprintf(stderr, " -coarse \tcoarsest rate (%g)\n", coarse);
    std::fprintf(stderr, " -conft  \tconfidence decay time (%g)\n", conft);
    std::fprintf(stderr, " -confa  \tconfidence asymptotic value (%g)\n", confa);
    std::fprintf(stderr, " -lambda \tlambda (%g)\n", lambda);
    std::fprintf(stderr, " -init   \tinitialization type [poisson/black/avg/none] (%s)\n", init.c_str());
    std::fprintf(stderr, " -psigma \tGaussian patch weights (%g)\n", psigma);
    std::fprintf(stderr, " -showpyr\tPREFIX write intermediate pyramid results\n");
    std::fprintf(stderr, " -shownnf\tFILENAME write illustration of the final NNF\n");
}

int run(int argc, char **argv)
{
    // main.cpp:68-82, in the same order
    const int patch_side = std::atoi(pick_option(&argc, &argv, "patch", "7"));
    int iterations = std::atoi(pick_option(&argc, &argv, "iters", "200"));
    const std::string method = pick_option(&argc, &argv, "method", "nlmeans");
    int scales = std::atoi(pick_option(&argc, &argv, "scales", "11"));
    float coarse = (float)std::atof(pick_option(&argc, &argv, "coarse", "0.1"));
    const float conft = (float)std::atof(pick_option(&argc, &argv, "conft", "5.0"));
    const float confa = (float)std::atof(pick_option(&argc, &argv, "confa", "0.1"));
    const float lambda = (float)std::atof(pick_option(&argc, &argv, "lambda", "0.05"));
    const std::string init_name = pick_option(&argc, &argv, "init", "poisson");
    const float psigma = (float)std::atof(pick_option(&argc, &argv, "psigma", "10000.0"));
    const std::string show_nnf = pick_option(&argc, &argv, "shownnf", "");
    const std::string show_pyramid = pick_option(&argc, &argv, "showpyr", "");
    const std::string depth_name = pick_option(&argc, &argv, "depth", "");
    const float lambda_rgb = (float)std::atof(pick_option(&argc, &argv, "lambdar", "0.05"));
    const float lambda_d = (float)std::atof(pick_option(&argc, &argv, "lambdad", "0.05"));
    const std::string seed_text = pick_option(&argc, &argv, "seed", "");
    const bool dry_run = pick_option(&argc, &argv, "dryrun", nullptr) != nullptr;
    if (argc < 4) {
        usage(argv[0], method, patch_side, iterations, scales, coarse, conft, confa, lambda, init_name, psigma);
        return 1;
    }
    const std::string input_name = argv[1], mask_name = argv[2], output_name = argv[3];

    // the GPU context is made on first use: -dryrun needs none (unless -coarse 0)
    std::unique_ptr<Context> ctx_holder;
    auto gpu = [&]() -> Context & {
        if (!ctx_holder) ctx_holder.reset(new Context(0));
        return *ctx_holder;
    };
    Image<float> input = IOUtility::read_rgb_image(input_name);   // :117
    if (!dry_run) input = IOUtility::rgb_to_lab(gpu(), input);
    Image<float> mask_image = IOUtility::read_mono_image(mask_name);
    if (input.get_size() != mask_image.get_size())
        throw std::runtime_error("ERROR: Input image and mask should be of the same size");
    bool empty = true;   // :125-138
    Mask mask(mask_image.get_size_x(), mask_image.get_size_y());
    for (unsigned y = 0; y < mask_image.get_size_y(); y++)
        for (unsigned x = 0; x < mask_image.get_size_x(); x++)
            if (mask_image(x, y) > 0) {
                mask.mask(x, y);
                empty = false;
            }
    if (empty) throw std::runtime_error("ERROR: Mask should not be empty (all black)");
    const bool depth_exists = !depth_name.empty();   // :140-147
    if (depth_exists) input = IOUtility::cat(input, IOUtility::read_mono_image(depth_name));

    if (coarse == 0) {   // :166-178: the patch no smaller than twice the coarsest hole
        const int W = (int)mask.get_size_x(), H = (int)mask.get_size_y();
        std::vector<float> dt((size_t)W * H);
        gpu().check(nli_distance_transform(gpu().get(), mask.raw(), W, H, dt.data()), "distance transform");
        float max_dist = 0;
        for (float v : dt) max_dist = (v > max_dist) ? v : max_dist;
        coarse = std::min(1.5f * patch_side / max_dist, 1.f);
    }
    // :180-198
    if (patch_side <= 0) throw std::runtime_error("ERROR: patch_side needs to be a positive integer.");
    if (coarse > 1 || coarse < 0) throw std::runtime_error("ERROR: coarsest rate should be between 0 and 1.");
    if (iterations < 0) throw std::runtime_error("ERROR: inpainting_iterations cannot be negative.");
    if (conft <= 0.f) throw std::runtime_error("ERROR: confidence_decay_time needs to be positive.");
    if (lambda > 1.f || lambda < 0.f) throw std::runtime_error("ERROR: lambda should be between 0 and 1.");
    if (scales < 1) throw std::runtime_error("ERROR: number of scales must be at least 1.");
    if (std::abs(coarse - 1.0f) < 0.00001f) {   // :200-203
        scales = 1;
        std::printf("\tnote: since the coarsest scale size ratio is 1.0, the number of scales is set to 1\n");
    }
    ImageInpainting::InitType init;   // :205-217
    if (init_name == "poisson") init = ImageInpainting::InitPoisson;
    else if (init_name == "black") init = ImageInpainting::InitBlack;
    else if (init_name == "avg") init = ImageInpainting::InitAvg;
    else if (init_name == "none") init = ImageInpainting::InitNone;
    else throw std::runtime_error("ERROR: Unknown initialization type");

    // :220-266: tolerance 0.1, 10 PatchMatch sweeps with 30 shots, CG 1e-6 / 1000
    const float tolerance = 0.1f;
    const Shape patch((unsigned)patch_side, (unsigned)patch_side);
    std::unique_ptr<AImageUpdating> updating;
    std::unique_ptr<APatchDistance> distance;
    if (method != "nlmeans" && method != "nlmedians" && method != "nlpoisson" && method != "nldepth")
        throw std::runtime_error("ERROR: Unknown method name");
    if (method == "nldepth" && !depth_exists) throw std::runtime_error("Specify depth input using -depth [filename]");
    // :287-303: the output file name carries the parameters (stream formatting as there)
    std::stringstream name;
    name << output_name << "_" << method;
    if (method == "nlpoisson") name << "_l" << lambda;
    name << "_sc" << scales << "_" << coarse;
    if (init == ImageInpainting::InitAvg) name << "_initavg";
    else if (init == ImageInpainting::InitBlack) name << "_init000";
    else name << "_initnone";
    name << "_ps" << patch_side << "_" << psigma;
    name << "_conf" << conft << "_" << confa;
    name << ".png";
    if (dry_run) {
        std::printf("%s\n", name.str().c_str());
        return 0;
    }
    Context &ctx = gpu();
    if (method == "nlmeans") {
        updating.reset(new HipPatchNonLocalMeans(ctx, patch, psigma));
        distance.reset(new L2NormPatchDistance(patch, psigma));
    } else if (method == "nlmedians") {
        updating.reset(new HipPatchNonLocalMedians(ctx, patch, psigma));
        distance.reset(new L1NormPatchDistance(patch, psigma));
    } else if (method == "nlpoisson") {
        updating.reset(new HipPatchNonLocalPoisson(ctx, patch, psigma, lambda, 0.000001f, 1000));
        distance.reset(new L2CombinedPatchDistance(lambda, patch, psigma));
    } else {
        updating.reset(new HipPatchNonLocalDepth(ctx, patch, psigma, lambda_rgb, lambda_d, 0.000001f, 1000));
        distance.reset(new L2RGBDPatchDistance(lambda_rgb, lambda_d, patch, psigma));
    }

    const uint64_t seed = seed_text.empty() ? (uint64_t)std::time(nullptr) : std::strtoull(seed_text.c_str(), nullptr, 10);
    HipPatchMatch patch_match(ctx, distance.get(), 10, 30, -1, seed);
    ImageInpainting inpainting(ctx, iterations, tolerance, scales,
                               ImageInpainting::calculate_subsampling_rate(coarse, (unsigned)scales), conft, confa,
                               init);
    inpainting.set_weights_updating(&patch_match);
    inpainting.set_image_updating(updating.get());
    inpainting.keep_intermediate(!show_nnf.empty() || !show_pyramid.empty());
    Image<float> output = inpainting.process(input, mask);

    IOUtility::write_rgb_image(name.str(), IOUtility::lab_to_rgb(ctx, output));   // :306-307
    if (depth_exists) IOUtility::write_mono_image("inpainted_depth.png", IOUtility::get_depth(output));

    if (!show_pyramid.empty()) {   // :313-323: every level of the output pyramid
        const std::vector<Image<float>> pyr = inpainting.get_output_image_pyramid();
        for (size_t n = 0; n < pyr.size(); n++) {
            std::stringstream level;
            level << show_pyramid << n << ".png";
            IOUtility::write_rgb_image(level.str(), IOUtility::lab_to_rgb(ctx, pyr[n]));
        }
    }
    if (!show_nnf.empty()) {   // :326-372: hole boundary and NNF discontinuities in black
        const std::vector<Image<Point>> nnfs = inpainting.get_nnf_pyramid();
        const Image<Point> last = nnfs.back();
        Image<float> show = IOUtility::lab_to_rgb(ctx, output);
        for (unsigned c = 0; c < show.get_number_of_channels(); c++)
            for (unsigned y = 1; y < show.get_size_y(); y++)
                for (unsigned x = 1; x < show.get_size_x(); x++) {
                    const bool m = mask.test((int)x, (int)y);
                    if (m != mask.test((int)x - 1, (int)y) || m != mask.test((int)x, (int)y - 1)) show(x, y, c) = 0;
                    if (m) {
                        // the offset nnf(p) - p differs from the left or upper neighbour's
                        const Point a = last(x - 1, y), b = last(x, y), u = last(x, y - 1);
                        if (a.x - (int)(x - 1) != b.x - (int)x || a.y - (int)y != b.y - (int)y) show(x, y, c) = 0;
                        if (u.x - (int)x != b.x - (int)x || u.y - (int)(y - 1) != b.y - (int)y) show(x, y, c) = 0;
                    }
                }
        IOUtility::write_rgb_image(show_nnf, show);
    }
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    try {
        return run(argc, argv);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
}
