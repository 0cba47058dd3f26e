/* nli_plugins.hpp -- C++ plugin classes over the libnli C-ABI (include/nli.h).
 *
 * The reference (Tom-Zheng/non-local-inpainting) composes its inpainting from three
 * plugin interfaces: APatchDistance (a_patch_distance.h:24-49), PatchMatch
 * (patch_match.h:30-80) and AImageUpdating (a_image_updating.h:24-52), wired into
 * ImageInpainting (image_inpainting.h:30-120) by main.cpp:245-284.  This header gives
 * the same classes, with the same names, constructor arguments and call semantics,
 * computing on the GPU through the C-ABI:
 *
 *   HipPatchMatch                 : PatchMatch      (nli_patchmatch)
 *   HipPatchNonLocalMeans         : AImageUpdating  (nli_update, NL-means)
 *   HipPatchNonLocalMedians       : AImageUpdating  (nli_update, NL-medians)
 *   HipPatchNonLocalPoisson       : AImageUpdating  (nli_update, NL-Poisson + CG)
 *   HipPatchNonLocalDepth         : AImageUpdating  (nli_update, two NL-Poisson solves)
 *   ImageInpainting               the multiscale driver of image_inpainting.cpp:68-557,
 *                                 with every numerical step a C-ABI call
 *   IOUtility                     rgb_to_lab / lab_to_rgb on the GPU (nli_rgb_to_lab,
 *                                 nli_lab_to_rgb), cat / separate / get_depth on the host,
 *                                 PNG files (read/write_rgb_image, read/write_mono_image),
 *                                 PGM files (read/write_pgm_image), mask_to_greyscale,
 *                                 probability_to_greyscale
 *
 * Differences from the reference interface, all deliberate:
 *   - PatchMatch::calculate is virtual and PatchMatch is abstract: the GPU subclass is
 *     the implementation (there is no CPU PatchMatch here to fall back to).
 *   - APatchDistance::describe() replaces the per-pair virtual calculate(Point, Point),
 *     which cannot be a GPU boundary: it names the distance (kind, lambda(s), patch
 *     side, sigma) so the PatchMatch kernel evaluates it for whole fields.
 *   - Every GPU object takes an nli::Context (one device context per host thread).
 *   - HipPatchMatch draws its random numbers from a counter-based generator seeded per
 *     call (seed, call index) instead of glibc rand(); runs are reproducible.
 *
 * Containers follow image.h/mask.h: Image<T> copies share the buffer (so update()
 * changes the caller's image, as in the reference), Image<T>(FixedImage<T>) deep-copies,
 * pixels are HWC with index C*(W*y+x)+c (image.hpp:481-485), masks hold one byte per
 * pixel, NNFs are Image<Point> with (-1,-1) = no entry.
 *
 * Errors from the library are thrown as std::runtime_error carrying nli_last_error().
 * HipPatchMatch::calculate returns an empty field on inconsistent sizes, exactly like
 * patch_match.cpp:71-76.
 */
#ifndef NLI_PLUGINS_HPP_
#define NLI_PLUGINS_HPP_

#include "nli.h"
#include "nli_io.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace nli {

/* ---- geometry (point.h, shape.h) ---- */
struct Point {
    int x = 0, y = 0;
    Point() = default;
    Point(int x_, int y_) : x(x_), y(y_) {}
    bool operator==(const Point &o) const { return x == o.x && y == o.y; }
    bool operator!=(const Point &o) const { return !(*this == o); }
};
static_assert(sizeof(Point) == 2 * sizeof(int32_t), "Point must be two packed int32 (NNF layout)");

struct Shape {
    unsigned size_x = 0, size_y = 0;
    Shape() = default;
    Shape(unsigned sx, unsigned sy) : size_x(sx), size_y(sy) {}
    bool operator==(const Shape &o) const { return size_x == o.size_x && size_y == o.size_y; }
    bool operator!=(const Shape &o) const { return !(*this == o); }
    bool is_empty() const { return size_x == 0 || size_y == 0; }
};

/* ---- device context ---- */
class Context {
public:
    explicit Context(int device = 0)
    {
        if (nli_ctx_create(device, &h_) != NLI_OK || !h_)
            throw std::runtime_error("nli_ctx_create failed: no usable GPU");
    }
    /* wraps a context created elsewhere (not destroyed here) */
    explicit Context(nli_ctx *borrowed) : h_(borrowed), owned_(false)
    {
        if (!h_) throw std::invalid_argument("null nli_ctx");
    }
    ~Context()
    {
        if (owned_) nli_ctx_destroy(h_);
    }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    nli_ctx *get() const { return h_; }
    /* throws on a non-zero return code; EmptySourceError for NLI_E_EMPTY_SOURCE */
    void check(int rc, const char *what) const;

private:
    nli_ctx *h_ = nullptr;
    bool owned_ = true;
};

struct EmptySourceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void Context::check(int rc, const char *what) const
{
    if (rc == NLI_OK) return;
    std::string msg = std::string(what) + ": " + nli_last_error(h_);
    if (rc == NLI_E_EMPTY_SOURCE) throw EmptySourceError(msg);
    throw std::runtime_error(msg);
}

/* ---- images (image.h:38-174) ---- */
template <class T>
class Image;

template <class T>
class FixedImage {
public:
    FixedImage() = default;
    FixedImage(unsigned size_x, unsigned size_y, unsigned channels = 1, const T &value = T())
    {
        init(size_x, size_y, channels, value);
    }
    FixedImage(Shape size, unsigned channels = 1, const T &value = T()) { init(size.size_x, size.size_y, channels, value); }
    virtual ~FixedImage() = default;

    bool is_empty() const { return !data_ || data_->empty(); }
    bool is_not_empty() const { return !is_empty(); }
    unsigned get_size_x() const { return sx_; }
    unsigned get_size_y() const { return sy_; }
    Shape get_size() const { return Shape(sx_, sy_); }
    unsigned get_number_of_channels() const { return nc_; }
    const T &operator()(unsigned x, unsigned y, unsigned c = 0) const { return (*data_)[get_index(x, y, c)]; }
    const T &operator()(Point p, unsigned c = 0) const { return (*this)(p.x, p.y, c); }
    const T &at(unsigned x, unsigned y, unsigned c = 0) const
    {
        if (x >= sx_ || y >= sy_ || c >= nc_) throw std::out_of_range("image index out of range");
        return (*this)(x, y, c);
    }
    const T *raw() const { return data_ ? data_->data() : nullptr; }
    size_t elements() const { return data_ ? data_->size() : 0; }
    FixedImage<T> clone() const
    {
        FixedImage<T> c;
        c.copy_from(*this);
        return c;
    }

protected:
    std::shared_ptr<std::vector<T>> data_;
    unsigned sx_ = 0, sy_ = 0, nc_ = 0;

    void init(unsigned sx, unsigned sy, unsigned nc, const T &v)
    {
        sx_ = sx;
        sy_ = sy;
        nc_ = nc;
        data_ = std::make_shared<std::vector<T>>((size_t)sx * sy * nc, v);
    }
    void copy_from(const FixedImage<T> &o)
    {
        sx_ = o.sx_;
        sy_ = o.sy_;
        nc_ = o.nc_;
        data_ = o.data_ ? std::make_shared<std::vector<T>>(*o.data_) : nullptr;
    }
    size_t get_index(unsigned x, unsigned y, unsigned c) const { return (size_t)nc_ * ((size_t)sx_ * y + x) + c; }
};

template <class T>
class Image : public FixedImage<T> {
public:
    Image() = default;
    Image(unsigned size_x, unsigned size_y, unsigned channels = 1, const T &value = T())
        : FixedImage<T>(size_x, size_y, channels, value) {}
    Image(Shape size, unsigned channels = 1, const T &value = T()) : FixedImage<T>(size, channels, value) {}
    Image(const Image<T> &o) = default;              // shares the buffer
    Image<T> &operator=(const Image<T> &o) = default;  // shares the buffer
    Image(const FixedImage<T> &o) { this->copy_from(o); }   // deep copy (image.h:139)
    Image<T> &operator=(const FixedImage<T> &o)              // deep copy (image.h:143)
    {
        this->copy_from(o);
        return *this;
    }

    using FixedImage<T>::operator();
    T &operator()(unsigned x, unsigned y, unsigned c = 0) { return (*this->data_)[this->get_index(x, y, c)]; }
    T &operator()(Point p, unsigned c = 0) { return (*this)(p.x, p.y, c); }
    T *raw() { return this->data_ ? this->data_->data() : nullptr; }
    const T *raw() const { return FixedImage<T>::raw(); }
    void fill(const T &v)
    {
        if (this->data_) std::fill(this->data_->begin(), this->data_->end(), v);
    }
    Image<T> clone() const
    {
        Image<T> c;
        c.copy_from(*this);
        return c;
    }
};

/* ---- masks (mask.h:36-183); one byte per pixel, 1 = in the set ---- */
class Mask;

class FixedMask : public FixedImage<uint8_t> {
public:
    FixedMask() = default;
    FixedMask(unsigned size_x, unsigned size_y, bool value = false) : FixedImage<uint8_t>(size_x, size_y, 1, value) {}
    FixedMask(Shape size, bool value = false) : FixedImage<uint8_t>(size, 1, value) {}
    /* out-of-range points test false (mask.cpp:223-230) */
    bool test(int x, int y) const
    {
        return x >= 0 && y >= 0 && (unsigned)x < sx_ && (unsigned)y < sy_ && (*data_)[(size_t)sx_ * y + x] != 0;
    }
    bool test(Point p) const { return test(p.x, p.y); }
    /* row-major point list (mask.cpp:322-329) */
    std::vector<Point> get_masked_points() const
    {
        std::vector<Point> pts;
        for (unsigned y = 0; y < sy_; y++)
            for (unsigned x = 0; x < sx_; x++)
                if ((*data_)[(size_t)sx_ * y + x]) pts.emplace_back((int)x, (int)y);
        return pts;
    }
};

class Mask : public FixedMask {
public:
    Mask() = default;
    Mask(unsigned size_x, unsigned size_y, bool value = false) : FixedMask(size_x, size_y, value) {}
    Mask(Shape size, bool value = false) : FixedMask(size, value) {}
    void mask(unsigned x, unsigned y) { (*data_)[(size_t)sx_ * y + x] = 1; }
    void unmask(unsigned x, unsigned y) { (*data_)[(size_t)sx_ * y + x] = 0; }
    uint8_t *raw() { return data_ ? data_->data() : nullptr; }
    const uint8_t *raw() const { return FixedMask::raw(); }
    Mask clone() const
    {
        Mask c;
        c.copy_from(*this);
        return c;
    }
};

/* ---- IOUtility: colour conversion and RGB-D channels (io_utility.h:51-65) ----
 * rgb_to_lab / lab_to_rgb run on the GPU (nli_rgb_to_lab / nli_lab_to_rgb) and take the
 * Context every GPU object takes; images they do not convert come back as deep copies,
 * as the reference returns its argument.  cat / separate / get_depth are the reference's
 * host-side channel copies with its messages.  The image files (read_rgb_image,
 * read_mono_image, write_rgb_image, write_mono_image) are PNG only, through nli_io.h in
 * libnli_plugins.so; where the reference prints a message and exits, these throw
 * std::runtime_error with the message. */
struct IOUtility {
    static void io_check(int rc, const char *prefix = "")
    {
        if (rc != NLI_OK) throw std::runtime_error(std::string(prefix) + nli_io_last_error());
    }
    static Image<float> read_rgb_image(const std::string &name)   // io_utility.cpp:143-168
    {
        int W = 0, H = 0;
        io_check(nli_image_info(name.c_str(), &W, &H, nullptr, nullptr), "read_rgb_image: ");
        Image<float> img((unsigned)W, (unsigned)H, 3u);
        io_check(nli_read_rgb_image(name.c_str(), img.raw(), (long)W * H * 3));
        return img;
    }
    static Image<float> read_mono_image(const std::string &name)   // io_utility.cpp:119-141
    {
        int W = 0, H = 0;
        io_check(nli_image_info(name.c_str(), &W, &H, nullptr, nullptr), "read_mono_image: ");
        Image<float> img((unsigned)W, (unsigned)H, 1u);
        io_check(nli_read_mono_image(name.c_str(), img.raw(), (long)W * H));
        return img;
    }
    static void write_rgb_image(const std::string &name, FixedImage<float> image)   // io_utility.cpp:171-190
    {
        io_check(nli_write_rgb_image(name.c_str(), image.raw(), (int)image.get_size_x(), (int)image.get_size_y(),
                                     (int)image.get_number_of_channels()));
    }
    static void write_mono_image(const std::string &name, FixedImage<float> image)   // io_utility.cpp:192-220
    {
        io_check(nli_write_mono_image(name.c_str(), image.raw(), (int)image.get_size_x(), (int)image.get_size_y(),
                                      (int)image.get_number_of_channels()));
    }

    // PGM without iio (io_utility.cpp:26-117): an empty image where the reference
    // returns one (a file it cannot read); a bad ASCII sample, where the reference exits,
    // throws.
    static Image<float> read_pgm_image(const std::string &name)
    {
        int W = 0, H = 0;
        if (nli_pgm_info(name.c_str(), &W, &H) != NLI_OK) return Image<float>();
        Image<float> img((unsigned)W, (unsigned)H, 1u);
        io_check(nli_read_pgm_image(name.c_str(), img.raw(), (long)W * H), "read_pgm_image: ");
        return img;
    }
    static void write_pgm_image(const std::string &name, FixedImage<float> image)
    {
        io_check(nli_write_pgm_image(name.c_str(), image.raw(), (int)image.get_size_x(), (int)image.get_size_y(),
                                     (int)image.get_number_of_channels()));
    }
    // Masked points white (255), the rest black (io_utility.cpp:223-236).
    static Image<float> mask_to_greyscale(FixedMask mask)
    {
        Image<float> out(mask.get_size(), 1u, 0.0f);
        for (const Point &p : mask.get_masked_points()) out((unsigned)p.x, (unsigned)p.y) = 255.0f;
        return out;
    }
    // [0, 1] to [0, 255], clamped in double (io_utility.cpp:243-256).
    static Image<float> probability_to_greyscale(FixedImage<float> probabilities)
    {
        Image<float> out(probabilities.get_size(), 1u, 0.0f);
        for (unsigned y = 0; y < probabilities.get_size_y(); y++)
            for (unsigned x = 0; x < probabilities.get_size_x(); x++)
                out(x, y) = (float)(255.0 * std::fmax(0.0, std::fmin(1.0, (double)probabilities(x, y))));
        return out;
    }
    static Image<float> rgb_to_lab(Context &ctx, FixedImage<float> image)   // io_utility.cpp:268-289
    {
        if (image.get_number_of_channels() != 3) return Image<float>(image);
        Image<float> out(image.get_size(), 3u);
        ctx.check(nli_rgb_to_lab(ctx.get(), image.raw(), (int)image.get_size_x(), (int)image.get_size_y(), 3,
                                 out.raw()),
                  "rgb_to_lab");
        return out;
    }
    static Image<float> lab_to_rgb(Context &ctx, FixedImage<float> image)   // io_utility.cpp:292-318
    {
        const unsigned C = image.get_number_of_channels();
        if (C != 3 && C != 4) return Image<float>(image);
        Image<float> out(image.get_size(), 3u);
        ctx.check(nli_lab_to_rgb(ctx.get(), image.raw(), (int)image.get_size_x(), (int)image.get_size_y(), (int)C,
                                 out.raw()),
                  "lab_to_rgb");
        return out;
    }
    static Image<float> cat(FixedImage<float> image, FixedImage<float> depth)   // io_utility.cpp:453-470
    {
        // the reference throws only when both channel counts are wrong (`&&`, :455)
        if (image.get_number_of_channels() != 3 && depth.get_number_of_channels() != 1)
            throw std::runtime_error("cat: Channels do not match.");
        if (image.get_size() != depth.get_size()) throw std::runtime_error("cat: Size does not match.");
        Image<float> out(image.get_size(), 4u);
        for (unsigned y = 0; y < image.get_size_y(); y++)
            for (unsigned x = 0; x < image.get_size_x(); x++) {
                for (unsigned c = 0; c < 3; c++) out(x, y, c) = image.at(x, y, c);
                out(x, y, 3) = depth.at(x, y);
            }
        return out;
    }
    static void separate(FixedImage<float> input, Image<float> &rgb, Image<float> &depth)   // :491-504
    {
        if (input.get_number_of_channels() != 4) throw std::runtime_error("Separate: Channels do not match.");
        if (rgb.get_size() != input.get_size() || rgb.get_number_of_channels() < 3 ||
            depth.get_size() != input.get_size())
            throw std::out_of_range("image index out of range");
        for (unsigned y = 0; y < input.get_size_y(); y++)
            for (unsigned x = 0; x < input.get_size_x(); x++) {
                for (unsigned c = 0; c < 3; c++) rgb(x, y, c) = input(x, y, c);
                depth(x, y) = input(x, y, 3);
            }
    }
    static Image<float> get_depth(FixedImage<float> input)   // io_utility.cpp:506-513
    {
        if (input.get_number_of_channels() != 4) throw std::runtime_error("Separate: Channels do not match.");
        Image<float> rgb(input.get_size(), 3u), depth(input.get_size());
        separate(input, rgb, depth);
        return depth;
    }
};

/* ---- patch distances (a_patch_distance.h, l2_*_patch_distance.h, l1_*) ---- */
class APatchDistance {
public:
    APatchDistance(Shape patch_size, float gaussian_sigma) : patch_size_(patch_size), sigma_(gaussian_sigma) {}
    virtual ~APatchDistance() = default;
    /* the distance as the PatchMatch kernel evaluates it */
    virtual nli_dist_params describe() const = 0;
    Shape get_patch_size() const { return patch_size_; }
    void set_patch_size(Shape s) { patch_size_ = s; }
    float get_gaussian_sigma() const { return sigma_; }
    void set_gaussian_sigma(float s) { sigma_ = s; }

protected:
    Shape patch_size_;
    float sigma_;
    nli_dist_params base(int kind) const
    {
        if (patch_size_.size_x != patch_size_.size_y)
            throw std::invalid_argument("patch distances use square patches (size_x == size_y)");
        nli_dist_params d{};
        d.kind = kind;
        d.patch = (int)patch_size_.size_x;
        d.sigma = sigma_;
        return d;
    }
};

/* l2_norm_patch_distance.cpp:22-104 */
class L2NormPatchDistance : public APatchDistance {
public:
    L2NormPatchDistance(Shape patch_size, float gaussian_sigma) : APatchDistance(patch_size, gaussian_sigma) {}
    nli_dist_params describe() const override { return base(NLI_DIST_L2); }
};

/* l1_norm_patch_distance.cpp:22-105 */
class L1NormPatchDistance : public APatchDistance {
public:
    L1NormPatchDistance(Shape patch_size, float gaussian_sigma) : APatchDistance(patch_size, gaussian_sigma) {}
    nli_dist_params describe() const override { return base(NLI_DIST_L1); }
};

/* l2_combined_patch_distance.cpp:29-148 */
class L2CombinedPatchDistance : public APatchDistance {
public:
    L2CombinedPatchDistance(float lambda, Shape patch_size, float gaussian_sigma)
        : APatchDistance(patch_size, gaussian_sigma), lambda_(lambda) {}
    nli_dist_params describe() const override
    {
        nli_dist_params d = base(NLI_DIST_L2COMBINED);
        d.lambda = lambda_;
        return d;
    }

private:
    float lambda_;
};

/* l2_rgbd_patch_distance.cpp:31-117 (four channels) */
class L2RGBDPatchDistance : public APatchDistance {
public:
    L2RGBDPatchDistance(float lambda_rgb, float lambda_d, Shape patch_size, float gaussian_sigma)
        : APatchDistance(patch_size, gaussian_sigma), lambda_rgb_(lambda_rgb), lambda_d_(lambda_d) {}
    nli_dist_params describe() const override
    {
        nli_dist_params d = base(NLI_DIST_L2RGBD);
        d.lambda = lambda_rgb_;
        d.lambda_d = lambda_d_;
        return d;
    }

private:
    float lambda_rgb_, lambda_d_;
};

/* ---- PatchMatch (patch_match.h:30-80) ---- */
class PatchMatch {
public:
    PatchMatch(APatchDistance *distance_calculation, int iteration_count, int random_shots_limit = 20,
               int search_window_size = -1)
        : distance_(distance_calculation), iterations_(iteration_count), shots_(random_shots_limit),
          window_(search_window_size) {}
    virtual ~PatchMatch() = default;
    /* NNF from every target-mask point to a source-mask point; initial_field empty =
     * random initialisation (patch_match.cpp:57-302) */
    virtual Image<Point> calculate(FixedImage<float> source, FixedMask source_mask, FixedImage<float> target,
                                   FixedMask target_mask, Image<Point> initial_field = Image<Point>()) = 0;
    int get_iteration_count() const { return iterations_; }
    void set_iteration_count(int n) { iterations_ = n; }
    int get_search_window_size() const { return window_; }
    void set_search_window_size(int n) { window_ = n; }
    int get_random_shots_limit() const { return shots_; }
    void set_random_shots_limit(int n) { shots_ = n; }
    void set_distance_calculation(APatchDistance *d) { distance_ = d; }
    APatchDistance *get_distance_calculation() const { return distance_; }

protected:
    APatchDistance *distance_;
    int iterations_, shots_, window_;
};

class HipPatchMatch : public PatchMatch {
public:
    HipPatchMatch(Context &ctx, APatchDistance *distance_calculation, int iteration_count,
                  int random_shots_limit = 20, int search_window_size = -1, uint64_t seed = 1)
        : PatchMatch(distance_calculation, iteration_count, random_shots_limit, search_window_size), ctx_(ctx),
          seed_(seed) {}

    Image<Point> calculate(FixedImage<float> source, FixedMask source_mask, FixedImage<float> target,
                           FixedMask target_mask, Image<Point> initial_field = Image<Point>()) override
    {
        // patch_match.cpp:71-76: inconsistent sizes give an empty field
        if ((!initial_field.is_empty() && initial_field.get_size() != target.get_size()) ||
            source.get_size() != source_mask.get_size() || target.get_size() != target_mask.get_size() ||
            !distance_)
            return Image<Point>();
        if (source.get_number_of_channels() != target.get_number_of_channels())
            throw std::invalid_argument("PatchMatch: source and target channels differ");
        const nli_dist_params d = distance_->describe();
        nli_pm_params pm{};
        pm.iterations = iterations_;
        pm.shots = shots_;
        pm.window = window_;
        pm.first_step = 0;
        const Shape ts = target.get_size();
        Image<Point> out(ts, 1, Point(-1, -1));
        distances_ = Image<float>(ts, 1, 0.0f);
        const uint64_t call_seed = seed_ * 0x9E3779B97F4A7C15ull + calls_++;
        ctx_.check(nli_patchmatch(ctx_.get(), &d, &pm, source.raw(), source_mask.raw(), (int)source.get_size_x(),
                                  (int)source.get_size_y(), target.raw(), target_mask.raw(), (int)ts.size_x,
                                  (int)ts.size_y, (int)target.get_number_of_channels(),
                                  initial_field.is_empty() ? nullptr
                                                           : reinterpret_cast<const int32_t *>(initial_field.raw()),
                                  call_seed, reinterpret_cast<int32_t *>(out.raw()), distances_.raw()),
                   "PatchMatch::calculate");
        return out;
    }
    /* distance of every target point to its match in the last calculate() (FLT_MAX
     * elsewhere) */
    FixedImage<float> get_last_distances() const { return distances_; }
    void set_seed(uint64_t seed)
    {
        seed_ = seed;
        calls_ = 0;
    }

private:
    Context &ctx_;
    uint64_t seed_, calls_ = 0;
    Image<float> distances_;
};

/* ---- image updating (a_image_updating.h:24-52) ---- */
class AImageUpdating {
public:
    AImageUpdating(Shape patch_size, float gaussian_sigma) : patch_size_(patch_size), sigma_(gaussian_sigma) {}
    virtual ~AImageUpdating() = default;
    /* updates `image` on the inpainting domain (in place, through the shared buffer) and
     * returns the sum of squared changes */
    virtual double update(Image<float> image, Image<float> orig_image, FixedMask inpainting_domain,
                          FixedMask extended_inpainting_domain, FixedImage<Point> nnf,
                          FixedImage<float> confidence_mask) = 0;
    float get_gaussian_sigma() const { return sigma_; }
    void set_gaussian_sigma(float s) { sigma_ = s; }
    Shape get_patch_size() const { return patch_size_; }
    void set_patch_size(Shape s) { patch_size_ = s; }

protected:
    Shape patch_size_;
    float sigma_;
};

/* shared body of the GPU updaters: one nli_update call */
class HipImageUpdating : public AImageUpdating {
public:
    HipImageUpdating(Context &ctx, int kind, Shape patch_size, float gaussian_sigma)
        : AImageUpdating(patch_size, gaussian_sigma), ctx_(ctx), kind_(kind) {}

    double update(Image<float> image, Image<float> orig_image, FixedMask inpainting_domain,
                  FixedMask extended_inpainting_domain, FixedImage<Point> nnf,
                  FixedImage<float> confidence_mask) override
    {
        const Shape s = image.get_size();
        if (orig_image.get_size() != s || inpainting_domain.get_size() != s ||
            extended_inpainting_domain.get_size() != s || nnf.get_size() != s ||
            (!confidence_mask.is_empty() && confidence_mask.get_size() != s))
            throw std::invalid_argument("update: image, masks, NNF and confidence must have the same size");
        if (patch_size_.size_x != patch_size_.size_y)
            throw std::invalid_argument("updates use square patches (size_x == size_y)");
        nli_upd_params u = params();
        double total = 0.0;
        nli_cg_stats st{};
        ctx_.check(nli_update(ctx_.get(), &u, image.raw(), orig_image.raw(), inpainting_domain.raw(),
                              extended_inpainting_domain.raw(), reinterpret_cast<const int32_t *>(nnf.raw()),
                              confidence_mask.is_empty() ? nullptr : confidence_mask.raw(), (int)s.size_x,
                              (int)s.size_y, (int)image.get_number_of_channels(), &total, &st),
                   "AImageUpdating::update");
        last_stats_ = st;
        return total;
    }
    /* CG statistics of the last update (Poisson and depth updaters) */
    const nli_cg_stats &get_last_cg_stats() const { return last_stats_; }

protected:
    Context &ctx_;
    int kind_;
    nli_cg_stats last_stats_{};
    virtual nli_upd_params params() const
    {
        nli_upd_params u{};
        u.kind = kind_;
        u.patch = (int)patch_size_.size_x;
        u.sigma = sigma_;
        return u;
    }
};

/* patch_non_local_means.cpp:23-100 */
class HipPatchNonLocalMeans : public HipImageUpdating {
public:
    HipPatchNonLocalMeans(Context &ctx, Shape patch_size, float gaussian_sigma)
        : HipImageUpdating(ctx, NLI_UPD_MEANS, patch_size, gaussian_sigma) {}
};

/* patch_non_local_medians.cpp:23-126 */
class HipPatchNonLocalMedians : public HipImageUpdating {
public:
    HipPatchNonLocalMedians(Context &ctx, Shape patch_size, float gaussian_sigma)
        : HipImageUpdating(ctx, NLI_UPD_MEDIANS, patch_size, gaussian_sigma) {}
};

/* patch_non_local_poisson.cpp:36-477; the constructor defaults of
 * patch_non_local_poisson.h:29-33.  precision NLI_CG_MIXED runs the CG in FP32 with FP64
 * dot products and an FP64 residual every refine_every iterations. */
class HipPatchNonLocalPoisson : public HipImageUpdating {
public:
    HipPatchNonLocalPoisson(Context &ctx, Shape patch_size, float gaussian_sigma, float lambda,
                            float conjugate_gradient_tolerance = 1e-10f,
                            int conjugate_gradient_iterations_limit = 1000, int precision = NLI_CG_FP64,
                            int refine_every = 0)
        : HipImageUpdating(ctx, NLI_UPD_POISSON, patch_size, gaussian_sigma), lambda_(lambda),
          tol_(conjugate_gradient_tolerance), limit_(conjugate_gradient_iterations_limit), precision_(precision),
          refine_(refine_every) {}

protected:
    float lambda_, tol_;
    int limit_, precision_, refine_;
    nli_upd_params params() const override
    {
        nli_upd_params u = HipImageUpdating::params();
        u.lambda = lambda_;
        u.cg_tol = tol_;
        u.cg_limit = limit_;
        u.cg_precision = precision_;
        u.cg_refine_every = refine_;
        return u;
    }
};

/* patch_non_local_depth.cpp:24-55: NL-Poisson on channels 0-2 (lambda_rgb) and on
 * channel 3 (lambda_d), both from `image` (the reference never reads orig_image here) */
class HipPatchNonLocalDepth : public HipPatchNonLocalPoisson {
public:
    HipPatchNonLocalDepth(Context &ctx, Shape patch_size, float gaussian_sigma, float lambda_rgb, float lambda_d,
                          float conjugate_gradient_tolerance = 1e-10f, int conjugate_gradient_iterations_limit = 1000,
                          int precision = NLI_CG_FP64, int refine_every = 0)
        : HipPatchNonLocalPoisson(ctx, patch_size, gaussian_sigma, lambda_rgb, conjugate_gradient_tolerance,
                                  conjugate_gradient_iterations_limit, precision, refine_every),
          lambda_d_(lambda_d)
    {
        kind_ = NLI_UPD_DEPTH;
    }

protected:
    float lambda_d_;
    nli_upd_params params() const override
    {
        nli_upd_params u = HipPatchNonLocalPoisson::params();
        u.lambda_d = lambda_d_;
        return u;
    }
};

/* ---- the multiscale driver (image_inpainting.h:30-120, image_inpainting.cpp:68-557) ----
 * Every step calls the C-ABI with host buffers: the plain drop-in, one library call per
 * reference call.  The device-resident form of the same loop is nli_inpaint_dev. */
class ImageInpainting {
public:
    enum InitType { InitBlack, InitAvg, InitNone, InitPoisson };

    ImageInpainting(Context &ctx, int iterations_amount, float tolerance, int scales_amount, float subsampling_rate,
                    float confidence_decay_time, float confidence_asymptotic_value, InitType initialization_type)
        : ctx_(ctx), iterations_(iterations_amount), tolerance_(tolerance), scales_(scales_amount),
          rate_(subsampling_rate), decay_(confidence_decay_time), asym_(confidence_asymptotic_value),
          init_(initialization_type) {}

    /* image_inpainting.cpp:197-204 */
    static float calculate_subsampling_rate(float size_ratio, unsigned scales_amount)
    {
        return nli_subsampling_rate(size_ratio, (int)scales_amount);
    }

    void set_weights_updating(PatchMatch *patch_match) { patch_match_ = patch_match; }
    PatchMatch *get_weights_updating() const { return patch_match_; }
    void set_image_updating(AImageUpdating *image_updating) { image_updating_ = image_updating; }
    AImageUpdating *get_image_updating() const { return image_updating_; }
    int get_iterations_amount() const { return iterations_; }
    void set_iterations_amount(int n) { iterations_ = n; }
    float get_tolerance() const { return tolerance_; }
    void set_tolerance(float t) { tolerance_ = t; }
    int get_scales_amount() const { return scales_; }
    void set_scales_amount(int n) { scales_ = n; }
    float get_subsampling_rate() const { return rate_; }
    void set_subsampling_rate(float r) { rate_ = r; }
    float get_confidence_decay_time() const { return decay_; }
    void set_confidence_decay_time(float v) { decay_ = v; }
    float get_confidence_asymptotic_value() const { return asym_; }
    void set_confidence_asymptotic_value(float v) { asym_ = v; }
    /* inpainting iterations run at each scale of the last process() (index 0 = finest) */
    const std::vector<int> &get_iterations_per_scale() const { return iterations_per_scale_; }

    /* image_inpainting.cpp:303-337: keep the input image pyramid and the NNF of every scale
     * (the output image pyramid and the mask pyramid are kept anyway) */
    void keep_intermediate(bool value) { keep_ = value; }
    std::vector<Image<float>> get_input_image_pyramid() const { return input_pyramid_; }
    std::vector<Image<float>> get_output_image_pyramid() const { return image_pyramid_; }
    std::vector<Mask> get_mask_pyramid() const { return mask_pyramid_; }
    /* the final NNF of each scale, coarsest first (image_inpainting.cpp:409-412) */
    std::vector<Image<Point>> get_nnf_pyramid() const { return nnf_pyramid_; }

    /* image_inpainting.cpp:68-190 */
    Image<float> process(FixedImage<float> in, FixedMask mask)
    {
        if (!patch_match_ || !image_updating_) throw std::runtime_error("ImageInpainting: plugins not set");
        if (scales_ < 1) throw std::invalid_argument("ERROR: number of scales must be at least 1.");
        if (in.get_size() != mask.get_size()) throw std::invalid_argument("image and mask sizes differ");
        const int C = (int)in.get_number_of_channels();
        std::vector<Image<float>> imgs((size_t)scales_);
        std::vector<Mask> masks((size_t)scales_);
        imgs[0] = Image<float>(in);   // deep copy: the input is not changed
        masks[0] = Mask(mask.get_size());
        std::memcpy(masks[0].raw(), mask.raw(), mask.elements());
        for (int i = 1; i < scales_; i++) {
            const int pw = (int)imgs[i - 1].get_size_x(), ph = (int)imgs[i - 1].get_size_y();
            const int w = nli_sample_size(pw, rate_), h = nli_sample_size(ph, rate_);
            imgs[i] = Image<float>((unsigned)w, (unsigned)h, (unsigned)C);
            masks[i] = Mask((unsigned)w, (unsigned)h);
            if (init_ != InitNone) {
                ctx_.check(nli_downsample_with_mask(ctx_.get(), imgs[i - 1].raw(), masks[i - 1].raw(), pw, ph, C, rate_,
                                                    kMaskSamplingThreshold, 0, imgs[i].raw(), masks[i].raw()),
                           "downsample_with_mask");
            } else {
                ctx_.check(nli_downsample(ctx_.get(), imgs[i - 1].raw(), pw, ph, C, rate_, imgs[i].raw()), "downsample");
                ctx_.check(nli_downsample_mask(ctx_.get(), masks[i - 1].raw(), pw, ph, rate_, kMaskSamplingThreshold,
                                               masks[i].raw()),
                           "downsample(mask)");
            }
        }
        input_pyramid_.clear();
        nnf_pyramid_.clear();
        if (keep_)   // image_inpainting.cpp:109-115: before the coarsest level is initialised
            for (const Image<float> &im : imgs) input_pyramid_.push_back(im.clone());
        initialize_coarsest(imgs.back(), masks.back());
        iterations_per_scale_.assign((size_t)scales_, 0);
        Image<Point> nnf;
        for (int i = scales_ - 1; i >= 0; i--) {
            Image<float> conf = confidence(masks[i]);
            if (i < scales_ - 1) nnf = propagate_weights(imgs[i], masks[i], conf, imgs[i + 1], masks[i + 1]);
            iterations_per_scale_[(size_t)i] = inpaint_internal(imgs[i], masks[i], conf, nnf);
            if (keep_) nnf_pyramid_.push_back(nnf);
        }
        image_pyramid_ = imgs;
        mask_pyramid_ = masks;
        return imgs.front();
    }

private:
    static constexpr float kMaskSamplingThreshold = 0.4f;   // image_inpainting.cpp:22
    Context &ctx_;
    PatchMatch *patch_match_ = nullptr;
    AImageUpdating *image_updating_ = nullptr;
    int iterations_;
    float tolerance_;
    int scales_;
    float rate_, decay_, asym_;
    InitType init_;
    std::vector<int> iterations_per_scale_;
    bool keep_ = false;
    std::vector<Image<float>> input_pyramid_, image_pyramid_;
    std::vector<Mask> mask_pyramid_;
    std::vector<Image<Point>> nnf_pyramid_;

    unsigned patch_side() const { return image_updating_->get_patch_size().size_x; }

    /* image_inpainting.cpp:104-154 */
    void initialize_coarsest(Image<float> img, FixedMask m)
    {
        const size_t n = (size_t)img.get_size_x() * img.get_size_y();
        const unsigned C = img.get_number_of_channels();
        if (init_ == InitNone) return;
        if (init_ == InitPoisson) {
            ctx_.check(nli_poisson_fill(img.raw(), m.raw(), (int)img.get_size_x(), (int)img.get_size_y(), (int)C),
                       "poisson_fill");
            return;
        }
        std::vector<float> color(C, 0.0f);
        if (init_ == InitAvg)
            for (unsigned c = 0; c < C; c++) {
                float sum = 0.0f;
                int count = 0;
                for (size_t k = 0; k < n; k++)
                    if (!m.raw()[k]) {
                        sum += img.raw()[k * C + c];
                        count++;
                    }
                if (count > 0) color[c] = sum / (float)count;
            }
        for (size_t k = 0; k < n; k++)
            if (m.raw()[k])
                for (unsigned c = 0; c < C; c++) img.raw()[k * C + c] = color[c];
    }

    /* image_inpainting.cpp:611-630 */
    Image<float> confidence(const FixedMask &m)
    {
        Image<float> conf(m.get_size(), 1, 1.0f);
        ctx_.check(nli_confidence(ctx_.get(), m.raw(), (int)m.get_size_x(), (int)m.get_size_y(), decay_, asym_,
                                  conf.raw()),
                   "confidence");
        return conf;
    }

    struct ScaleMasks {
        Mask ext, target, source;
    };
    /* extended domain, target and source of a scale (image_inpainting.cpp:359-384 with
     * erode, :496-502 and :543-548 without) */
    ScaleMasks scale_masks(const FixedMask &hole, bool erode)
    {
        ScaleMasks s{Mask(hole.get_size()), Mask(hole.get_size()), Mask(hole.get_size())};
        ctx_.check(nli_scale_masks(ctx_.get(), hole.raw(), (int)hole.get_size_x(), (int)hole.get_size_y(),
                                   (int)patch_side(), erode ? 1 : 0, s.ext.raw(), s.target.raw(), s.source.raw()),
                   "scale masks");
        return s;
    }

    /* image_inpainting.cpp:349-417; returns the iterations run */
    int inpaint_internal(Image<float> image, FixedMask domain, FixedImage<float> conf, Image<Point> &nnf)
    {
        Image<float> original = image.clone();
        ScaleMasks m = scale_masks(domain, true);   // throws EmptySourceError (:382-384)
        double total = std::numeric_limits<double>::max();
        int i = 0;
        for (i = 0; i < iterations_ && total > (double)tolerance_; i++) {
            nnf = patch_match_->calculate(image, m.source, image, m.target, nnf);
            total = image_updating_->update(image, original, domain, m.ext, nnf, conf);
        }
        return i;
    }

    /* image_inpainting.cpp:492-557 */
    Image<Point> propagate_weights(Image<float> upper, FixedMask upper_domain, FixedImage<float> upper_conf,
                                   FixedImage<float> lower, FixedMask lower_domain)
    {
        ScaleMasks lm = scale_masks(lower_domain, false);
        Image<Point> lnnf = patch_match_->calculate(lower, lm.source, lower, lm.target);
        if (lnnf.is_empty()) throw std::runtime_error("PatchMatch returned an empty field");
        const unsigned W = upper.get_size_x(), H = upper.get_size_y(), C = upper.get_number_of_channels();
        Image<Point> scaled(upper.get_size(), 1, Point(-1, -1));
        ctx_.check(nli_scale_nnf(ctx_.get(), reinterpret_cast<const int32_t *>(lnnf.raw()), lower_domain.raw(),
                                 (int)lower.get_size_x(), (int)lower.get_size_y(), (int)W, (int)H,
                                 reinterpret_cast<int32_t *>(scaled.raw())),
                   "scale NNF");
        Image<float> up(upper.get_size(), C);
        ctx_.check(nli_upsample(ctx_.get(), lower.raw(), (int)lower.get_size_x(), (int)lower.get_size_y(), (int)C,
                                (int)W, (int)H, up.raw()),
                   "upsample");
        const size_t n = (size_t)W * H;
        for (size_t k = 0; k < n; k++)   // :533-538
            if (upper_domain.raw()[k])
                for (unsigned c = 0; c < C; c++) upper.raw()[k * C + c] = up.raw()[k * C + c];
        ScaleMasks um = scale_masks(upper_domain, false);
        Image<Point> refined = patch_match_->calculate(up, um.source, up, um.target, scaled);
        image_updating_->update(upper, upper, upper_domain, um.ext, refined, upper_conf);
        return refined;
    }
};

}  // namespace nli

/* The reference's main.cpp:220-284 composition (distance/updater pair chosen by
 * params->method, PatchMatch with params->pm_*, ImageInpainting with the remaining
 * parameters) built from the classes above and run on host buffers: the plain drop-in,
 * one C-ABI call per reference call.  Exported by libnli_plugins.so. */
extern "C" int nli_plain_inpaint(nli_ctx *ctx, const nli_inpaint_params *params, const float *in,
                                 const uint8_t *mask, int W, int H, int C, float *out,
                                 nli_inpaint_stats *stats);
/* main.cpp:115-309 around that composition, without file I/O and argument parsing:
 * rgb (W*H*3, 0-255) -> IOUtility::rgb_to_lab, the mask taken where positive
 * (main.cpp:125-138), depth (W*H, optional) appended by IOUtility::cat, the parameter
 * rules of main.cpp:166-203 (params->coarse = 0: the rate from the largest distance to
 * the hole boundary; a rate within 1e-5 of 1: one scale), process, then
 * IOUtility::lab_to_rgb into rgb_out (W*H*3) and IOUtility::get_depth into depth_out
 * (W*H, with depth).  Exported by libnli_plugins.so; its results equal nli_inpaint_rgb's. */
extern "C" int nli_plain_inpaint_rgb(nli_ctx *ctx, const nli_inpaint_params *params, const float *rgb,
                                     const float *depth, const uint8_t *mask, int W, int H, float *rgb_out,
                                     float *depth_out, nli_inpaint_stats *stats);
/* message of the last failed nli_plain_inpaint(_rgb) on this host thread */
extern "C" const char *nli_plain_last_error(void);

#endif /* NLI_PLUGINS_HPP_ */
