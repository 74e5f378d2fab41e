// Frame encoder orchestration (frame_encoder.cu:88-352 in the reference):
// one stream, buffers allocated once per configuration, every stage on the
// device, and exactly two host synchronizations per frame -- one to learn the
// file size, one after copying the file out.
#include "encoder.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <memory>
#include <mutex>

#include "kernels.h"
#include "prefix.h"

namespace cjx {
namespace {

// Host bit writer over a growable word vector.
struct HostBits {
    std::vector<std::uint32_t> words = std::vector<std::uint32_t>(16, 0u);
    std::uint32_t pos{0};
    void write(std::uint32_t n, std::uint32_t v) {
        while ((pos + n + 64) / 32 >= words.size()) {
            words.resize(words.size() * 2, 0u);
        }
        BitSink s{words.data(), pos, static_cast<std::uint32_t>(words.size() * 32)};
        s.write(n, v);
        pos = s.pos;
    }
    void pad() { pos = (pos + 7u) & ~7u; }
    void append(const HostBits& o) { copy_bits(*this, o.words.data(), o.pos); }
};

constexpr Distr kDimEnc[4]{{false, 1, 9}, {false, 1, 13}, {false, 1, 18}, {false, 1, 30}};
constexpr Distr kGlobalScaleEnc[4]{{false, 1, 11}, {false, 2049, 11}, {false, 4097, 12}, {false, 8193, 16}};
constexpr Distr kQuantDcEnc[4]{{true, 16, 0}, {false, 1, 5}, {false, 1, 8}, {false, 1, 16}};
constexpr Distr kOrderEnc[4]{{true, 0x5F, 0}, {true, 0x13, 0}, {true, 0, 0}, {false, 0, 13}};

// write_prefix_histograms (histogram_writer.cc:59-83) for constant alphabets.
void prefix_histograms(HostBits& w, const std::uint32_t* hist, int length, int num_contexts, std::uint32_t se,
                       std::uint32_t msb, std::uint32_t lsb, std::uint8_t* depth, std::uint16_t* bits) {
    auto scratch{std::make_unique<HuffScratch>()};
    w.write(1, 0);
    if (num_contexts > 1) {
        w.write(1, 1);
        w.write(2, 0);
    }
    w.write(1, 1);
    write_uint_config(w, se, msb, lsb);
    write_var_len_u16(w, static_cast<std::uint32_t>(length - 1));
    std::fill(depth, depth + length, 0);
    std::fill(bits, bits + length, 0);
    if (length > 1) {
        build_and_store_huffman(hist, length, depth, bits, w, *scratch);
    }
}

// context_map.cc:43-168: the smallest of complex, complex+MTF and simple.
void complex_cmap(HostBits& w, const std::uint8_t* map, int n, bool mtf) {
    std::vector<std::uint8_t> sym(map, map + n);
    if (mtf) {
        std::uint8_t s[256];
        for (int i{0}; i < 256; ++i) {
            s[i] = static_cast<std::uint8_t>(i);
        }
        for (int i{0}; i < n; ++i) {
            int r{0};
            while (s[r] != map[i]) {
                ++r;
            }
            sym[i] = static_cast<std::uint8_t>(r);
            const std::uint8_t v{s[r]};
            for (; r > 0; --r) {
                s[r] = s[r - 1];
            }
            s[0] = v;
        }
    }
    std::uint32_t hist[16]{};
    for (std::uint8_t v : sym) {
        std::uint32_t t, nb, b;
        hybrid_uint(2, 0, 1, v, t, nb, b);
        ++hist[t];
    }
    int alpha{1};
    for (int i{0}; i < 16; ++i) {
        if (hist[i]) {
            alpha = i + 1;
        }
    }
    w.write(1, 0);
    w.write(1, mtf ? 1 : 0);
    std::uint8_t depth[16];
    std::uint16_t bits[16];
    prefix_histograms(w, hist, alpha, 1, 2, 0, 1, depth, bits);
    for (std::uint8_t v : sym) {
        std::uint32_t t, nb, b;
        hybrid_uint(2, 0, 1, v, t, nb, b);
        w.write(depth[t], bits[t]);
        w.write(nb, b);
    }
}

HostBits best_context_map(const std::uint8_t* map, int n, int nclusters) {
    HostBits cx, cm, sp;
    complex_cmap(cx, map, n, false);
    complex_cmap(cm, map, n, true);
    const HostBits* sel{&cx};
    if (cm.pos < sel->pos) {
        sel = &cm;
    }
    if (nclusters <= 8) {
        const std::uint32_t bpe{ceil_log2(static_cast<std::uint32_t>(nclusters))};
        sp.write(1, 1);
        sp.write(2, bpe);
        for (int i{0}; i < n; ++i) {
            sp.write(bpe, map[i]);
        }
        if (sp.pos <= sel->pos) {
            sel = &sp;
        }
    }
    return *sel;
}

// ac_context.h:108-131 on the host (the fixed 32-cluster context map).
int cluster_of(std::uint32_t ctx) {
    constexpr std::uint32_t count_contexts{15u * 37u};
    if (ctx < count_contexts) {
        const std::uint32_t bucket{ctx / 15u};
        const int group{static_cast<int>(ctx % 15u) < 7 ? 0 : 1};
        return group * 4 + (bucket < 4 ? 0 : (bucket < 12 ? 1 : (bucket < 24 ? 2 : 3)));
    }
    const std::uint32_t rel{ctx - count_contexts};
    const int group{static_cast<int>(rel / 458u) < 7 ? 0 : 1};
    const std::uint32_t z{rel % 458u};
    static const std::uint32_t edges[11]{32, 64, 96, 128, 160, 200, 240, 280, 328, 376, 424};
    int band{11};
    for (int i{0}; i < 11; ++i) {
        if (z < edges[i]) {
            band = i;
            break;
        }
    }
    return 8 + group * 12 + band;
}

const HostBits& ac_context_map_bits() {
    static const HostBits bits{[] {
        std::vector<std::uint8_t> map(kNumAcContexts);
        for (int i{0}; i < kNumAcContexts; ++i) {
            map[i] = static_cast<std::uint8_t>(cluster_of(static_cast<std::uint32_t>(i)));
        }
        return best_context_map(map.data(), kNumAcContexts, kNumClusters);
    }()};
    return bits;
}

// frame_assembly.cc:70-101 up to (not including) the data alphabet size:
// GroupHeader, single-leaf tree with `predictor`, the data histogram's fixed
// leading fields.
void modular_prefix(HostBits& w, int predictor) {
    w.write(1, 0);
    w.write(1, 1);
    w.write(2, 0);
    const std::uint32_t tokens[5]{0, static_cast<std::uint32_t>(predictor), 0, 0, 0};
    std::uint32_t th[16]{};
    int tl{1};
    for (std::uint32_t t : tokens) {
        ++th[t];
        tl = std::max(tl, static_cast<int>(t) + 1);
    }
    std::uint8_t td[16];
    std::uint16_t tb[16];
    prefix_histograms(w, th, tl, 6, 4, 2, 0, td, tb);
    for (std::uint32_t t : tokens) {
        w.write(td[t], tb[t]);
    }
    w.write(1, 0);  // data container: no LZ77
    w.write(1, 1);  // prefix codes
    write_uint_config(w, 4, 2, 0);
}

std::once_flag g_tables_once[64];
bool g_tables_ok[64];

// Test hook: the allocation that brings the countdown from 1 to 0 fails as
// if the device were out of memory (0: no failure pending).
std::atomic<int> g_fail_alloc_countdown{0};

bool injected_alloc_failure() {
    int n{g_fail_alloc_countdown.load()};
    while (n > 0) {
        if (g_fail_alloc_countdown.compare_exchange_weak(n, n - 1)) {
            return n == 1;
        }
    }
    return false;
}

}  // namespace

bool ensure_device_tables(int device) {
    std::call_once(g_tables_once[device & 63], [device] { g_tables_ok[device & 63] = init_device_tables(); });
    return g_tables_ok[device & 63];
}

void fail_allocation_after(int n) {
    g_fail_alloc_countdown.store(n < 0 ? 0 : n);
}

bool device_exists(int ordinal) {
    int count{0};
    if (hipGetDeviceCount(&count) != hipSuccess) {
        return false;
    }
    return ordinal >= 0 && ordinal < count;
}

bool device_has_matrix_cores(int ordinal) {
    hipDeviceProp_t p{};
    if (hipGetDeviceProperties(&p, ordinal) != hipSuccess) {
        return false;
    }
    const std::string arch{p.gcnArchName};
    return arch.rfind("gfx90a", 0) == 0 || arch.rfind("gfx94", 0) == 0 || arch.rfind("gfx95", 0) == 0;
}

std::size_t Encoder::max_output_size(std::uint32_t width, std::uint32_t height) {
    const std::size_t bw{width / 8u}, bh{height / 8u};
    const std::size_t nac{((bw + 31) / 32) * ((bh + 31) / 32)};
    const std::size_t ndc{((bw + 255) / 256) * ((bh + 255) / 256)};
    return 3 * bw * bh * 384 + 48 * bw * bh + (nac + ndc) * 2048 + 16384;
}

Encoder::~Encoder() {
    release();
}

void Encoder::release() {
    if (device_ >= 0) {
        (void)hipSetDevice(device_);
    }
    // A frame queued by encode_async and never finished may still run.
    for (hipStream_t s : {stream_, dc_stream_}) {
        if (s != nullptr) {
            (void)hipStreamSynchronize(s);
        }
    }
    pending_ = false;
    for (void* p : allocations_) {
        (void)hipFree(p);
    }
    allocations_.clear();
    input_ = nullptr;
    if (total_host_ != nullptr) {
        (void)hipHostFree(total_host_);
        total_host_ = nullptr;
    }
    if (stage_host_ != nullptr) {
        (void)hipHostFree(stage_host_);
        stage_host_ = nullptr;
    }
    for (hipEvent_t& e : stage_ready_) {
        if (e != nullptr) {
            (void)hipEventDestroy(e);
            e = nullptr;
        }
    }
    for (hipEvent_t& e : events_) {
        if (e != nullptr) {
            (void)hipEventDestroy(e);
            e = nullptr;
        }
    }
    for (hipEvent_t* e : {&ev_fork_, &ev_dc_done_, &ev_size_, &ev_end_, &ev_done_, &ev_caller_, &ev_input_free_}) {
        if (*e != nullptr) {
            (void)hipEventDestroy(*e);
            *e = nullptr;
        }
    }
    if (dc_stream_ != nullptr) {
        (void)hipStreamDestroy(dc_stream_);
        dc_stream_ = nullptr;
    }
    if (stream_ != nullptr) {
        (void)hipStreamDestroy(stream_);
        stream_ = nullptr;
    }
    null_buffers();
    total_dev_ = nullptr;
    configured_ = false;
    workspace_ = false;
}

// Every device buffer member back to null (their memory was freed with
// allocations_), so nothing can reach freed memory after a failed configure.
void Encoder::null_buffers() {
    coeffs_ = nullptr;
    acs_ = nullptr;
    qf_ = nullptr;
    cfl_part_ = nullptr;
    ytox_ = ytob_ = nullptr;
    ac_ = nullptr;
    dc_ = nullptr;
    nz_grid_ = nullptr;
    nz_count_ = nullptr;
    hists_ = fb_count_ = nullptr;
    codes_ = constants_ = blobs_ = layout_ = nullptr;
    blob_words_ = nullptr;
    tokens_ = nullptr;
    item_tokens_ = nullptr;
    item_list_ = item_count_ = hist_partial_ = ac_staging_ = dc_staging_ = nullptr;
    ac_part_bits_ = ac_group_bits_ = chunk_bits_ = dc_group_bits_ = ac_base_ = dc_base_ = nullptr;
    ac_group_bytes_ = dc_group_bytes_ = nullptr;
    out_ = nullptr;
    out_cap_ = 0;
    features_ = nullptr;
    export_ = false;
    const_gs_ = const_qdc_ = 0;
}

bool Encoder::configure(std::uint32_t width, std::uint32_t height, float distance, int device) {
    if (width == 0 || height == 0 || width % 8 || height % 8) {
        return false;
    }
    const Geometry geo{Geometry::of(width, height)};
    if (geo.num_dc > 32) {
        return false;  // per-group tables are sized for up to 32 DC groups
    }
    const bool same_shape{configured_ && device == device_ && width == geo_.width && height == geo_.height};
    if (pending_) {
        return false;  // a frame queued by encode_async is still owed to finish()
    }
    distance_ = distance;
    prm_ = make_frame_params(distance);
    prm_.mode = mode_;
    if (same_shape) {
        return true;
    }
    release();
    device_ = device;
    geo_ = geo;
    if (!setup(device)) {
        // Leave nothing half-built: no buffers, no shape a later configure
        // could mistake for its own, and every encode refused until a
        // configure succeeds.
        release();
        device_ = -1;
        geo_ = Geometry{};
        return false;
    }
    configured_ = true;
    return true;
}

bool Encoder::setup(int device) {
    if (hipSetDevice(device) != hipSuccess) {
        return false;
    }
    std::call_once(g_tables_once[device & 63], [device] { g_tables_ok[device & 63] = init_device_tables(); });
    if (!g_tables_ok[device & 63]) {
        return false;
    }
    if (hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&dc_stream_, hipStreamNonBlocking) != hipSuccess) {
        return false;
    }
    for (hipEvent_t* e : {&ev_done_, &ev_caller_, &ev_input_free_}) {
        if (hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess) {
            return false;
        }
    }
    // Bound to kernels' completion at launch (timing-capable events).
    for (hipEvent_t* e : {&ev_fork_, &ev_size_, &ev_dc_done_, &ev_end_}) {
        if (hipEventCreate(e) != hipSuccess) {
            return false;
        }
    }
    for (hipEvent_t& e : events_) {
        if (hipEventCreate(&e) != hipSuccess) {
            return false;
        }
    }
    if (hipHostMalloc(reinterpret_cast<void**>(&total_host_), sizeof(std::uint64_t)) != hipSuccess ||
        hipHostGetDevicePointer(reinterpret_cast<void**>(&total_dev_), total_host_, 0) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&stage_host_), 2 * kStageChunk) != hipSuccess) {
        return false;
    }
    for (hipEvent_t& e : stage_ready_) {
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            return false;
        }
    }
    return alloc();
}

template <class Get>
void Encoder::carve(Get& get) {
    const std::size_t nblk{geo_.nblocks};
    const std::size_t tiles{static_cast<std::size_t>(geo_.cmw) * geo_.cmh};
    const std::size_t regions{static_cast<std::size_t>(geo_.rw) * geo_.rh};
    const std::size_t nh{static_cast<std::size_t>(kNumClusters) + 2 * geo_.num_dc};
    get(coeffs_, 3 * nblk * 64 * sizeof(std::uint16_t));
    get(acs_, nblk);
    get(qf_, nblk * sizeof(std::int32_t));
    get(cfl_part_, regions * 3 * sizeof(float));
    get(ytox_, tiles);
    get(ytob_, tiles);
    get(ac_, 3 * nblk * 64 * sizeof(std::int16_t));
    get(dc_, 3 * nblk * sizeof(std::int32_t));
    get(nz_grid_, 3 * nblk);
    get(nz_count_, 3 * nblk * sizeof(std::uint16_t));
    get(hists_, nh * kHistStride * sizeof(std::uint32_t));
    get(fb_count_, 64 * sizeof(std::uint32_t));
    get(codes_, sizeof_codeset());
    get(constants_, sizeof_constants());
    get(blobs_, sizeof_dc_blobs());
    get(blob_words_, static_cast<std::size_t>(geo_.num_dc) * blob_words_per_group() * sizeof(std::uint32_t));
    get(tokens_, 3 * nblk * 64 * sizeof(std::uint32_t));
    get(item_tokens_, 3 * nblk * sizeof(std::uint16_t));
    get(item_list_, 3 * nblk * sizeof(std::uint32_t));
    get(item_count_, 2 * sizeof(std::uint32_t));
    get(hist_partial_, ac_partial_words() * sizeof(std::uint32_t));
    get(ac_staging_, static_cast<std::size_t>(geo_.num_ac) * ac_group_slot_words() * sizeof(std::uint32_t));
    get(dc_staging_, static_cast<std::size_t>(geo_.num_dc) * dc_group_slot_words() * sizeof(std::uint32_t));
    get(ac_part_bits_, geo_.num_ac * ac_group_parts() * sizeof(std::uint64_t));
    get(ac_group_bytes_, geo_.num_ac * sizeof(std::uint32_t));
    get(ac_group_bits_, geo_.num_ac * sizeof(std::uint64_t));
    get(chunk_bits_, static_cast<std::size_t>(geo_.num_dc) * dc_chunks_per_group() * sizeof(std::uint64_t));
    get(dc_group_bytes_, geo_.num_dc * sizeof(std::uint32_t));
    get(dc_group_bits_, geo_.num_dc * sizeof(std::uint64_t));
    get(layout_, sizeof_layout());
    get(ac_base_, geo_.num_ac * sizeof(std::uint64_t));
    get(dc_base_, geo_.num_dc * sizeof(std::uint64_t));
}

bool Encoder::alloc() {
    out_cap_ = (max_output_size(geo_.width, geo_.height) + 64 + 3) & ~static_cast<std::size_t>(3);
    bool ok{true};
    auto get = [&](auto*& p, std::size_t bytes) {
        void* q{nullptr};
        if (ok && !injected_alloc_failure() && hipMalloc(&q, bytes < 256 ? 256 : bytes) == hipSuccess) {
            allocations_.push_back(q);
            p = static_cast<std::remove_reference_t<decltype(p)>>(q);
        } else {
            ok = false;
        }
    };
    carve(get);
    get(out_, out_cap_);
    const_gs_ = const_qdc_ = 0;
    return ok;
}

// Consecutive 256-byte aligned ranges from offset 0 of a workspace (`base`
// null: sizing only).
namespace {
struct Bump {
    std::uint8_t* base;
    std::size_t used{0};
    template <class T>
    void operator()(T*& p, std::size_t bytes) {
        p = base != nullptr ? reinterpret_cast<T*>(base + used) : nullptr;
        used += (std::max<std::size_t>(bytes, 1) + 255) & ~static_cast<std::size_t>(255);
    }
};
}  // namespace

std::size_t Encoder::workspace_bytes(std::uint32_t width, std::uint32_t height) {
    Encoder probe;
    probe.geo_ = Geometry::of(width, height);
    Bump b{nullptr};
    probe.carve(b);
    probe.null_buffers();
    return b.used;
}

bool Encoder::bind_workspace(std::uint32_t width, std::uint32_t height, float distance, int device, void* base,
                             std::size_t bytes) {
    if (!allocations_.empty() || stream_ != nullptr || base == nullptr || width == 0 || height == 0 ||
        width % 8 || height % 8) {
        return false;  // an encoder that owns allocations is never rebound
    }
    const Geometry geo{Geometry::of(width, height)};
    if (geo.num_dc > 32 || workspace_bytes(width, height) > bytes) {
        return false;
    }
    if (hipSetDevice(device) != hipSuccess) {
        return false;
    }
    std::call_once(g_tables_once[device & 63], [device] { g_tables_ok[device & 63] = init_device_tables(); });
    if (!g_tables_ok[device & 63]) {
        return false;
    }
    const bool same{workspace_ && configured_ && device == device_ && geo.width == geo_.width &&
                    geo.height == geo_.height &&
                    static_cast<void*>(coeffs_) == base};
    device_ = device;
    distance_ = distance;
    prm_ = make_frame_params(distance);
    prm_.mode = mode_;
    if (!same) {
        geo_ = geo;
        Bump b{static_cast<std::uint8_t*>(base)};
        carve(b);
        out_ = nullptr;
        out_cap_ = 0;
        const_gs_ = const_qdc_ = 0;  // the constants depend on the shape
    }
    workspace_ = true;
    configured_ = true;
    return true;
}

bool Encoder::fill_job(FrameJob& j, hipStream_t s, std::uint8_t* out_dev, std::size_t cap, std::uint64_t* size_dev) {
    if (!workspace_ || !configured_ || hipSetDevice(device_) != hipSuccess ||
        !build_constants(prm_.global_scale, prm_.quant_dc, s)) {
        return false;
    }
    j.geo = geo_;
    j.prm = prm_;
    j.coeffs = coeffs_;
    j.acs = acs_;
    j.qf = qf_;
    j.cfl_part = cfl_part_;
    j.ytox = ytox_;
    j.ytob = ytob_;
    j.ac = ac_;
    j.dc = dc_;
    j.nz_grid = nz_grid_;
    j.nz_count = nz_count_;
    j.hists = hists_;
    j.fb_count = fb_count_;
    j.codes = static_cast<CodeSet*>(codes_);
    j.k = static_cast<const Constants*>(constants_);
    j.blobs = static_cast<DcBlobs*>(blobs_);
    j.blob_words = blob_words_;
    j.tokens = tokens_;
    j.item_tokens = item_tokens_;
    j.item_list = item_list_;
    j.item_count = item_count_;
    j.hist_partial = hist_partial_;
    j.ac_staging = ac_staging_;
    j.dc_staging = dc_staging_;
    j.ac_part_bits = ac_part_bits_;
    j.ac_group_bytes = ac_group_bytes_;
    j.ac_group_bits = ac_group_bits_;
    j.chunk_bits = chunk_bits_;
    j.dc_group_bytes = dc_group_bytes_;
    j.dc_group_bits = dc_group_bits_;
    j.lay = static_cast<Layout*>(layout_);
    j.ac_base = ac_base_;
    j.dc_base = dc_base_;
    j.out = reinterpret_cast<std::uint32_t*>(out_dev);
    j.cap = cap;
    j.size_out = size_dev;
    j.clr = entropy_clears();
    return true;
}

#ifdef CJX_REGION_CUT
bool Encoder::region_only(const std::uint8_t* luma, std::size_t luma_pitch, const std::uint8_t* chroma,
                          std::size_t chroma_pitch) {
    if (!configured_ || workspace_ || pending_ || hipSetDevice(device_) != hipSuccess) {
        return false;
    }
    launch_region(stream_, geo_, prm_, luma, luma_pitch, chroma, chroma_pitch, coeffs_, acs_, qf_, cfl_part_,
                  item_count_, entropy_clears());
    return hipStreamSynchronize(stream_) == hipSuccess;
}
#endif

// One frame queued on `s` in dependency order (the two entropy chains one
// after the other); nothing waits on the host unless the constant bit
// strings change (a new shape or distance), which copies them in order on
// `s` and waits for that copy.
bool Encoder::encode_on(hipStream_t s, const std::uint8_t* luma, std::size_t luma_pitch, const std::uint8_t* chroma,
                        std::size_t chroma_pitch, std::uint8_t* out_dev, std::size_t cap, std::uint64_t* size_dev) {
    if (!workspace_ || !configured_ || hipSetDevice(device_) != hipSuccess ||
        !build_constants(prm_.global_scale, prm_.quant_dc, s)) {
        return false;
    }
    auto* codes{static_cast<CodeSet*>(codes_)};
    auto* k{static_cast<const Constants*>(constants_)};
    auto* blobs{static_cast<DcBlobs*>(blobs_)};
    auto* lay{static_cast<Layout*>(layout_)};
    launch_region(s, geo_, prm_, luma, luma_pitch, chroma, chroma_pitch, coeffs_, acs_, qf_, cfl_part_, item_count_,
                  entropy_clears());
    launch_tile_quantize(s, geo_, prm_, coeffs_, acs_, qf_, cfl_part_, ac_, nz_grid_, nz_count_,
                         ItemList{item_list_, item_count_});
    launch_dc_maps(s, geo_, prm_, coeffs_, acs_, cfl_part_, dc_, ytox_, ytob_);
    launch_dc_hist(s, geo_, dc_, acs_, qf_, ytox_, ytob_, hists_ + kNumClusters * kHistStride,
                   hists_ + (kNumClusters + geo_.num_dc) * kHistStride, fb_count_);
    launch_huffman(s, kNumClusters, static_cast<int>(2 * geo_.num_dc), hists_, codes);
    launch_dc_sizes(s, geo_, k, codes, fb_count_, blobs, blob_words_, dc_, acs_, qf_, ytox_, ytob_, chunk_bits_,
                    dc_group_bytes_, dc_group_bits_, dc_staging_);
    launch_dc_emit(s, geo_, dc_, acs_, qf_, ytox_, ytob_, codes, blobs, blob_words_, chunk_bits_, dc_staging_);
    launch_ac_tokens(s, geo_, ac_, acs_, nz_grid_, nz_count_, tokens_, item_tokens_, hist_partial_, hists_,
                     ItemList{item_list_, item_count_});
    launch_huffman(s, 0, kNumClusters, hists_, codes);
    launch_ac_group_emit(s, geo_, acs_, item_tokens_, tokens_, codes, ac_staging_, ac_part_bits_);
    launch_layout(s, geo_, k, codes, ac_part_bits_, dc_group_bytes_, dc_group_bits_, ac_group_bytes_,
                  ac_group_bits_, blobs, lay, ac_base_, dc_base_, size_dev);
    launch_assemble(s, geo_, k, lay, dc_staging_, ac_staging_, dc_group_bits_, ac_group_bits_, ac_part_bits_,
                    dc_base_, ac_base_, out_dev, cap);
    return hipGetLastError() == hipSuccess;
}

// Constant bit strings of a configuration (codestream.cc, frame_assembly.cc).
bool Encoder::build_constants(std::uint32_t global_scale, std::uint32_t quant_dc, hipStream_t s) {
    if (global_scale == const_gs_ && quant_dc == const_qdc_) {
        return true;
    }
    HostBits head;
    head.write(8, 0xff);
    head.write(8, 0x0a);
    const bool small{geo_.width % 8 == 0 && geo_.height % 8 == 0 && geo_.width <= 256 && geo_.height <= 256};
    head.write(1, small);
    if (small) {
        head.write(5, geo_.height / 8 - 1);
    } else {
        write_u32(head, kDimEnc, geo_.height);
    }
    head.write(3, 0);
    if (small) {
        head.write(5, geo_.width / 8 - 1);
    } else {
        write_u32(head, kDimEnc, geo_.width);
    }
    head.write(1, 1);  // ImageMetadata all_default
    head.write(1, 1);  // CustomTransformData all_default
    head.pad();
    // FrameHeader (codestream.cc:67-85).
    const Distr ft[4]{{true, 0, 0}, {true, 1, 0}, {true, 2, 0}, {true, 3, 0}};
    const Distr ups[4]{{true, 1, 0}, {true, 2, 0}, {true, 4, 0}, {true, 8, 0}};
    const Distr passes[4]{{true, 1, 0}, {true, 2, 0}, {true, 3, 0}, {false, 4, 3}};
    const Distr blend[4]{{true, 0, 0}, {true, 1, 0}, {true, 2, 0}, {false, 3, 2}};
    const Distr name[4]{{true, 0, 0}, {false, 0, 4}, {false, 16, 5}, {false, 48, 10}};
    head.write(1, 0);
    write_u32(head, ft, 0);
    head.write(1, 0);
    head.write(2, 0);  // flags = 0 (U64)
    write_u32(head, ups, 1);
    head.write(3, 2);
    head.write(3, 2);
    write_u32(head, passes, 1);
    head.write(1, 0);
    write_u32(head, blend, 0);
    head.write(1, 1);
    write_u32(head, name, 0);
    head.write(1, 1);
    head.write(2, 0);  // extensions = 0 (U64)

    HostBits dcg;  // vardct_frame.cc:533-540
    dcg.write(1, 1);
    write_u32(dcg, kGlobalScaleEnc, global_scale);
    write_u32(dcg, kQuantDcEnc, quant_dc);
    dcg.write(1, 1);
    dcg.write(1, 1);
    dcg.write(1, 0);

    HostBits acp;  // frame_assembly.cc:131-146 + histogram_writer.cc:85-104
    acp.write(1, 1);
    acp.write(ceil_log2(geo_.num_ac), 0);
    write_u32(acp, kOrderEnc, 0);
    acp.write(1, 0);
    acp.append(ac_context_map_bits());
    acp.write(1, 1);
    for (int c{0}; c < kNumClusters; ++c) {
        write_uint_config(acp, 4, 2, 0);
    }

    HostBits pre;  // DecodeVarDCTDC extra_precision + modular prefix
    pre.write(2, 0);
    modular_prefix(pre, 5);
    HostBits mid;
    modular_prefix(mid, 0);

    if (!constants_fit(head.pos, acp.pos, pre.pos, mid.pos)) {
        return false;
    }
    std::vector<std::uint8_t> host(sizeof_constants());
    fill_constants(host.data(), head.words.data(), head.pos, dcg.words.data(), dcg.pos, acp.words.data(),
                   acp.pos, pre.words.data(), pre.pos, mid.words.data(), mid.pos);
    const hipStream_t on{s != nullptr ? s : stream_};
    if (hipMemcpyAsync(constants_, host.data(), host.size(), hipMemcpyHostToDevice, on) != hipSuccess ||
        hipStreamSynchronize(on) != hipSuccess) {
        return false;
    }
    const_gs_ = global_scale;
    const_qdc_ = quant_dc;
    return true;
}

void Encoder::mark(Stage s, hipStream_t on) {
    if (profiling_ == 1 || (profiling_ == 2 && (s == Stage::kRegion || s == Stage::kQuantize))) {
        (void)hipEventRecord(events_[static_cast<int>(s)], on != nullptr ? on : stream_);
    }
}

// Device file -> caller's host buffer through the two pinned slots: chunk k
// goes to slot k % 2; the host copies a chunk out while the next one is in
// flight, then reuses its slot for chunk k + 2.
namespace {

// True for host memory the runtime can DMA into directly (allocated or
// registered through HIP); such an output skips the pinned staging copy.
bool host_pinned(const void* p) {
    hipPointerAttribute_t a{};
    if (p == nullptr || hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeHost;
}

}  // namespace

// The caller's output as a device address when it is page-locked host memory
// mapped for the device: the assembly kernel then writes the file there
// itself (over the host link, as a DMA would) and no copy follows it.
std::uint8_t* Encoder::mapped_output(std::uint8_t* dst) const {
    hipPointerAttribute_t a{};
    if (dst == nullptr || hipPointerGetAttributes(&a, dst) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    if (a.type != hipMemoryTypeHost || a.devicePointer == nullptr || a.hostPointer == nullptr) {
        return nullptr;
    }
    // Base-relative: correct whether the runtime reports the allocation's
    // base or the queried address in both fields.  The kernel stores words,
    // so an unaligned output takes the copy path instead.
    std::uint8_t* const dev{static_cast<std::uint8_t*>(a.devicePointer) +
                            (dst - static_cast<std::uint8_t*>(a.hostPointer))};
    return reinterpret_cast<std::uintptr_t>(dev) % 4 == 0 ? dev : nullptr;
}

// The readback's first half, queued behind the frame.  The layout kernel
// writes the size into the pinned word (its completion is `ev_size_`).  A mapped
// page-locked output `direct` holds the file once the assembly has finished;
// any other output (or none yet, for encode_async) gets the first two chunks
// staged in the pinned buffer (bytes past the end are never used).
bool Encoder::queue_readback(std::uint8_t* direct) {
    mark(Stage::kReadback, tail_);
    staged_ = 0;
    direct_dst_ = direct;
    if (direct == nullptr) {
        for (int slot{0}; slot < 2 && staged_ < out_cap_; ++slot) {
            if (!stage_chunk(slot, std::min(kStageChunk, out_cap_ - staged_))) {
                return false;
            }
        }
    }
    pending_ = true;
    return true;
}

bool Encoder::stage_chunk(int slot, std::size_t n) {
    stage_len_[slot] = n;
    const bool ok{hipMemcpyAsync(stage_host_ + slot * kStageChunk, reinterpret_cast<const std::uint8_t*>(out_) + staged_,
                                 n, hipMemcpyDeviceToHost, tail_) == hipSuccess &&
                  hipEventRecord(stage_ready_[slot], tail_) == hipSuccess};
    staged_ += n;
    return ok;
}

// Polls `e` until it has completed (false on an error).
bool Encoder::spin_on(hipEvent_t e) {
    for (;;) {
        const hipError_t r{hipEventQuery(e)};
        if (r == hipSuccess) {
            return true;
        }
        if (r != hipErrorNotReady) {
            return false;
        }
        __builtin_ia32_pause();
    }
}

// The end of a frame: the assembly kernel's own completion when it writes
// the output and ends the frame, else an event after the last copy.  The
// host polls (only tens of microseconds of work remain): polling saves the
// wake-up latency of a blocking wait, which is of the same order.
bool Encoder::wait_done() {
    if (end_bound_) {
        end_bound_ = false;
        return spin_on(ev_end_);
    }
    return hipEventRecord(ev_done_, tail_) == hipSuccess && spin_on(ev_done_);
}

// The readback's second half: waits for the size, then moves the bytes into
// `dst` -- one DMA when `dst` is pinned, else through the two staging chunks,
// each host copy overlapping the next chunk's DMA.
bool Encoder::finish(std::uint8_t* dst, std::size_t cap, std::size_t* size) {
    if (!pending_) {
        return false;
    }
    pending_ = false;
    // Block until the DC chain (off the critical path, done tens of
    // microseconds before the layout), then poll for the size, or for the
    // frame's end when no size event was bound (the assembly writes the
    // file into `dst`): a blocking wait on a late event would add its
    // wake-up latency, and the event's interrupt, to the frame.
    const bool ended{!size_bound_};
    if (hipEventSynchronize(ev_dc_done_) != hipSuccess || !(ended ? wait_done() : spin_on(ev_size_))) {
        return false;
    }
    const std::uint64_t total{*total_host_};
    *size = static_cast<std::size_t>(total);
    bool ok{total <= out_cap_};
    if (ok && total <= cap && dst != nullptr) {
        if (direct_dst_ != nullptr && dst == direct_dst_) {
            // The assembly kernel writes the file into `dst`; the frame's end
            // (waited for above without a size event, else below) completes it.
        } else if (host_pinned(dst)) {
            ok = hipMemcpyAsync(dst, out_, total, hipMemcpyDeviceToHost, tail_) == hipSuccess;
        } else {
            std::size_t done{0};
            for (int slot{0}; slot < 2 && staged_ < total && stage_len_[slot] == 0; ++slot) {
                ok = ok && stage_chunk(slot, std::min<std::size_t>(kStageChunk, total - staged_));
            }
            for (int slot{0}; ok && done < total; slot ^= 1) {
                if (hipEventSynchronize(stage_ready_[slot]) != hipSuccess) {
                    return false;
                }
                const std::size_t n{std::min<std::size_t>(stage_len_[slot], total - done)};
                std::memcpy(dst + done, stage_host_ + slot * kStageChunk, n);
                done += n;
                stage_len_[slot] = 0;
                if (staged_ < total) {
                    ok = stage_chunk(slot, std::min<std::size_t>(kStageChunk, total - staged_));
                }
            }
        }
    }
    const bool copied{direct_dst_ == nullptr || dst != direct_dst_};
    stage_len_[0] = stage_len_[1] = 0;
    direct_dst_ = nullptr;
    if (!ended || copied) {
        mark(Stage::kCount, tail_);
        if (!wait_done()) {
            return false;
        }
    }
    if (profiling_ != 0) {
        for (int i{0}; i < static_cast<int>(Stage::kCount); ++i) {
            float ms{0.0f};
            if (profiling_ == 1 || i == static_cast<int>(Stage::kRegion)) {
                (void)hipEventElapsedTime(&ms, events_[i], events_[i + 1]);
            }
            stage_ms_[i] = ms;
        }
    }
    return ok;
}

bool Encoder::submit_nv12(const std::uint8_t* luma, std::size_t luma_pitch, const std::uint8_t* chroma,
                          std::size_t chroma_pitch, hipStream_t caller, bool release_caller, std::uint8_t* dst,
                          std::size_t cap) {
    if (!configured_ || workspace_ || pending_ || hipSetDevice(device_) != hipSuccess ||
        !build_constants(prm_.global_scale, prm_.quant_dc)) {
        return false;
    }
    // The frame starts after the work already queued on `caller` (the null
    // stream for the synchronous entry points: a caller that filled the
    // planes there, as callers of the reference did, is ordered before the
    // region kernel reads them).  A null stream with nothing left to run
    // needs no event: that saves a marker and a barrier packet ahead of the
    // region kernel on every frame of a synchronous loop.
    const bool settled{caller == nullptr && hipStreamQuery(nullptr) == hipSuccess};
    if (!settled &&
        (hipEventRecord(ev_caller_, caller) != hipSuccess || hipStreamWaitEvent(stream_, ev_caller_, 0) != hipSuccess)) {
        return false;
    }
    // With profiling on, the region kernel's own start and end times open the
    // region and quantize stages.  Its end is the fork point of the DC chain
    // (it leaves the coefficients the DC values come from and the CfL
    // partials the colour maps come from, and clears the entropy stage's
    // accumulators).
    const bool timed{profiling_ != 0};
    const hipEvent_t region_end{timed ? events_[static_cast<int>(Stage::kQuantize)] : ev_fork_};
    launch_region(stream_, geo_, prm_, luma, luma_pitch, chroma, chroma_pitch, coeffs_, acs_, qf_, cfl_part_,
                  item_count_, entropy_clears(),
                  timed ? events_[static_cast<int>(Stage::kRegion)] : nullptr, region_end,
                  export_ ? features_ : nullptr);
    // The planes are read by the region kernel only: later work on the
    // caller's stream may overwrite them once it has finished.
    if (release_caller && (hipEventRecord(ev_input_free_, stream_) != hipSuccess ||
                           hipStreamWaitEvent(caller, ev_input_free_, 0) != hipSuccess)) {
        return false;
    }
    // The quantizer lists the non-empty items for the tokenizer.
    launch_tile_quantize(stream_, geo_, prm_, coeffs_, acs_, qf_, cfl_part_, ac_, nz_grid_, nz_count_,
                         ItemList{item_list_, item_count_});
    // A mapped page-locked output receives the file from the assembly kernel.
    std::uint8_t* const direct{mapped_output(dst)};
    return run_entropy(direct != nullptr ? direct : reinterpret_cast<std::uint8_t*>(out_),
                       direct != nullptr ? cap : out_cap_, region_end, true, true,
                       direct == nullptr || profiling_ != 0) &&
           queue_readback(direct != nullptr ? dst : nullptr);
}

bool Encoder::encode_nv12(const std::uint8_t* luma, std::size_t luma_pitch, const std::uint8_t* chroma,
                          std::size_t chroma_pitch, std::uint8_t* out, std::size_t cap, std::size_t* size) {
    if (!submit_nv12(luma, luma_pitch, chroma, chroma_pitch, nullptr, false, out, cap)) {
        pending_ = false;
        (void)hipStreamSynchronize(stream_);
        (void)hipStreamSynchronize(dc_stream_);
        return false;
    }
    return finish(out, cap, size);
}

bool Encoder::encode_dct8_uniform(const std::uint8_t* luma, std::size_t luma_pitch, const std::uint8_t* chroma,
                                  std::size_t chroma_pitch, std::uint8_t* out, std::size_t cap, std::size_t* size) {
    if (!configured_ || pending_ || hipSetDevice(device_) != hipSuccess || !build_constants(prm_.global_scale, prm_.quant_dc)) {
        return false;
    }
    if (hipEventRecord(ev_caller_, nullptr) != hipSuccess || hipStreamWaitEvent(stream_, ev_caller_, 0) != hipSuccess) {
        return false;
    }
    mark(Stage::kRegion);
    launch_dct8_uniform(stream_, geo_, prm_, luma, luma_pitch, chroma, chroma_pitch, ac_, dc_, acs_, qf_, ytox_,
                        ytob_);
    mark(Stage::kQuantize);
    launch_nz(stream_, geo_, ac_, acs_, nz_grid_, nz_count_, entropy_clears());
    if (!run_entropy(reinterpret_cast<std::uint8_t*>(out_), out_cap_, nullptr, false) || !queue_readback(nullptr)) {
        pending_ = false;
        (void)hipStreamSynchronize(stream_);
        (void)hipStreamSynchronize(dc_stream_);
        return false;
    }
    return finish(out, cap, size);
}

bool Encoder::encode_async(const std::uint8_t* luma, std::size_t luma_pitch, const std::uint8_t* chroma,
                           std::size_t chroma_pitch, hipStream_t caller) {
    const bool ok{submit_nv12(luma, luma_pitch, chroma, chroma_pitch, caller, true, nullptr, 0)};
    if (!ok && pending_) {
        pending_ = false;
        (void)hipStreamSynchronize(stream_);
        (void)hipStreamSynchronize(dc_stream_);
    }
    return ok;
}

// Host planes -> the encoder's device NV12 frame (width pitch), row blocks
// through the pinned staging chunks: the host copy of one chunk overlaps the
// DMA of the other.  Pinned planes are copied by one 2D DMA each.
bool Encoder::upload_nv12(const std::uint8_t* luma, std::size_t luma_pitch, const std::uint8_t* chroma,
                          std::size_t chroma_pitch) {
    if (!configured_ || pending_ || hipSetDevice(device_) != hipSuccess) {
        return false;
    }
    const std::size_t w{geo_.width}, h{geo_.height};
    if (input_ == nullptr) {
        void* p{nullptr};
        if (hipMalloc(&p, w * h * 3 / 2) != hipSuccess) {
            return false;
        }
        allocations_.push_back(p);
        input_ = static_cast<std::uint8_t*>(p);
    }
    struct Plane {
        const std::uint8_t* src;
        std::size_t pitch, rows;
        std::uint8_t* dst;
    };
    const Plane planes[2]{{luma, luma_pitch, h, input_}, {chroma, chroma_pitch, h / 2, input_ + w * h}};
    if (host_pinned(luma) && host_pinned(chroma)) {
        for (const Plane& pl : planes) {
            if (hipMemcpy2DAsync(pl.dst, w, pl.src, pl.pitch, w, pl.rows, hipMemcpyHostToDevice, stream_) !=
                hipSuccess) {
                return false;
            }
        }
        return hipStreamSynchronize(stream_) == hipSuccess;
    }
    const std::size_t rows_per_chunk{kStageChunk / w};
    if (rows_per_chunk == 0) {
        return false;
    }
    bool busy[2]{false, false};
    int slot{0};
    for (const Plane& pl : planes) {
        for (std::size_t r0{0}; r0 < pl.rows; r0 += rows_per_chunk) {
            const std::size_t nr{std::min(rows_per_chunk, pl.rows - r0)};
            if (busy[slot] && hipEventSynchronize(stage_ready_[slot]) != hipSuccess) {
                return false;
            }
            std::uint8_t* st{stage_host_ + slot * kStageChunk};
            for (std::size_t r{0}; r < nr; ++r) {
                std::memcpy(st + r * w, pl.src + (r0 + r) * pl.pitch, w);
            }
            if (hipMemcpyAsync(pl.dst + r0 * w, st, nr * w, hipMemcpyHostToDevice, stream_) != hipSuccess ||
                hipEventRecord(stage_ready_[slot], stream_) != hipSuccess) {
                return false;
            }
            busy[slot] = true;
            slot ^= 1;
        }
    }
    return hipStreamSynchronize(stream_) == hipSuccess;
}

bool Encoder::encode_host(const std::uint8_t* luma, std::size_t luma_pitch, const std::uint8_t* chroma,
                          std::size_t chroma_pitch, std::uint8_t* out, std::size_t cap, std::size_t* size) {
    return upload_nv12(luma, luma_pitch, chroma, chroma_pitch) &&
           encode_nv12(input_, geo_.width, input_ + static_cast<std::size_t>(geo_.width) * geo_.height, geo_.width,
                       out, cap, size);
}

bool Encoder::encode_coefficients(const std::int16_t* ac, const std::int32_t* dc, const std::int8_t* acs,
                                  const std::int32_t* qf, const std::int8_t* ytox, const std::int8_t* ytob,
                                  std::uint32_t global_scale, std::uint32_t quant_dc, std::uint8_t* out,
                                  std::size_t cap, std::size_t* size) {
    if (!configured_ || pending_ || hipSetDevice(device_) != hipSuccess || !build_constants(global_scale, quant_dc)) {
        return false;
    }
    const std::size_t nblk{geo_.nblocks};
    const std::size_t tiles{static_cast<std::size_t>(geo_.cmw) * geo_.cmh};
    std::vector<std::int8_t> zeros(tiles, 0);
    const bool ok{
        hipMemcpyAsync(ac_, ac, 3 * nblk * 64 * sizeof(std::int16_t), hipMemcpyHostToDevice, stream_) == hipSuccess &&
        hipMemcpyAsync(dc_, dc, 3 * nblk * sizeof(std::int32_t), hipMemcpyHostToDevice, stream_) == hipSuccess &&
        hipMemcpyAsync(acs_, acs, nblk, hipMemcpyHostToDevice, stream_) == hipSuccess &&
        hipMemcpyAsync(qf_, qf, nblk * sizeof(std::int32_t), hipMemcpyHostToDevice, stream_) == hipSuccess &&
        hipMemcpyAsync(ytox_, ytox != nullptr ? ytox : zeros.data(), tiles, hipMemcpyHostToDevice, stream_) ==
            hipSuccess &&
        hipMemcpyAsync(ytob_, ytob != nullptr ? ytob : zeros.data(), tiles, hipMemcpyHostToDevice, stream_) ==
            hipSuccess &&
        hipStreamSynchronize(stream_) == hipSuccess};
    if (!ok) {
        return false;
    }
    mark(Stage::kRegion);
    mark(Stage::kQuantize);
    launch_nz(stream_, geo_, ac_, acs_, nz_grid_, nz_count_, entropy_clears());
    if (!run_entropy(reinterpret_cast<std::uint8_t*>(out_), out_cap_, nullptr, false) || !queue_readback(nullptr)) {
        pending_ = false;
        (void)hipStreamSynchronize(stream_);
        (void)hipStreamSynchronize(dc_stream_);
        return false;
    }
    return finish(out, cap, size);
}

// The histograms (AC clusters, then DC and AcMetadata per DC group), the
// first-block counts and the DC header-blob table start every frame at zero.
Clears Encoder::entropy_clears() const {
    Clears c{};
    c.ptr[0] = hists_;
    c.words[0] = static_cast<std::uint32_t>((kNumClusters + 2 * geo_.num_dc) * kHistStride);
    c.ptr[1] = fb_count_;
    c.words[1] = 64;
    c.ptr[2] = static_cast<std::uint32_t*>(blobs_);
    c.words[2] = static_cast<std::uint32_t>(sizeof_dc_blobs() / 4);
    return c;
}

bool Encoder::run_entropy(std::uint8_t* out, std::size_t cap, hipEvent_t fork, bool listed, bool maps,
                          bool size_event) {
    auto* codes{static_cast<CodeSet*>(codes_)};
    auto* k{static_cast<const Constants*>(constants_)};
    auto* blobs{static_cast<DcBlobs*>(blobs_)};
    auto* lay{static_cast<Layout*>(layout_)};
    // The DC chain (histograms, codes, header blobs, chunk sizes, emission)
    // depends on nothing of the AC chain: it runs on the second stream from
    // `fork` (the stream's work so far when null) and joins before the layout.
    // With `maps` it first derives the DC values from the coefficients and
    // the tiles' colour maps from the CfL partials.  `size_event` binds ev_size_ to the layout kernel (the host
    // then reads the size before the frame's end).
    mark(Stage::kHistograms);
    const int ndc{static_cast<int>(2 * geo_.num_dc)};
    if (fork == nullptr) {
        fork = ev_fork_;
        if (hipEventRecord(fork, stream_) != hipSuccess) {
            return false;
        }
    }
    if (hipStreamWaitEvent(dc_stream_, fork, 0) != hipSuccess) {
        return false;
    }
    if (maps) {
        launch_dc_maps(dc_stream_, geo_, prm_, coeffs_, acs_, cfl_part_, dc_, ytox_, ytob_);
    }
    launch_dc_hist(dc_stream_, geo_, dc_, acs_, qf_, ytox_, ytob_, hists_ + kNumClusters * kHistStride,
                   hists_ + (kNumClusters + geo_.num_dc) * kHistStride, fb_count_);
    launch_huffman(dc_stream_, kNumClusters, ndc, hists_, codes);
    launch_dc_sizes(dc_stream_, geo_, k, codes, fb_count_, blobs, blob_words_, dc_, acs_, qf_, ytox_, ytob_,
                    chunk_bits_, dc_group_bytes_, dc_group_bits_, dc_staging_);
    // The DC chain's end is the emission kernel's own completion (the join
    // below and the host's first wait use it).
    launch_dc_emit(dc_stream_, geo_, dc_, acs_, qf_, ytox_, ytob_, codes, blobs, blob_words_, chunk_bits_,
                   dc_staging_, ev_dc_done_);
    launch_ac_tokens(stream_, geo_, ac_, acs_, nz_grid_, nz_count_, tokens_, item_tokens_, hist_partial_, hists_,
                     listed ? ItemList{item_list_, item_count_} : ItemList{nullptr, nullptr});
    mark(Stage::kCodes);
    launch_huffman(stream_, 0, kNumClusters, hists_, codes);
    mark(Stage::kSizes);
    launch_ac_group_emit(stream_, geo_, acs_, item_tokens_, tokens_, codes, ac_staging_, ac_part_bits_);
    // Join: the AC stream waits for the DC chain's end (which, forked at the
    // region kernel's end, is done well before the AC emission: the wait on
    // a completed event is the cheap one) and runs the layout, the assembly
    // and the readback.
    if (hipStreamWaitEvent(stream_, ev_dc_done_, 0) != hipSuccess) {
        return false;
    }
    tail_ = stream_;
    size_bound_ = size_event;
    // Without a size event the output is written by the assembly kernel and
    // nothing follows it: its own completion ends the frame.
    end_bound_ = !size_event;
    launch_layout(stream_, geo_, k, codes, ac_part_bits_, dc_group_bytes_, dc_group_bits_, ac_group_bytes_,
                  ac_group_bits_, blobs, lay, ac_base_, dc_base_, total_dev_, size_event ? ev_size_ : nullptr);
    mark(Stage::kEmit, tail_);
    launch_assemble(stream_, geo_, k, lay, dc_staging_, ac_staging_, dc_group_bits_, ac_group_bits_, ac_part_bits_,
                    dc_base_, ac_base_, out, cap, end_bound_ ? ev_end_ : nullptr);
    return hipGetLastError() == hipSuccess;
}

bool Encoder::export_features(bool on) {
    if (!configured_ || workspace_ || hipSetDevice(device_) != hipSuccess) {
        return false;
    }
    if (on && features_ == nullptr) {
        void* p{nullptr};
        if (hipMalloc(&p, static_cast<std::size_t>(geo_.rw) * geo_.rh * 108 * sizeof(float)) != hipSuccess) {
            return false;
        }
        allocations_.push_back(p);
        features_ = static_cast<float*>(p);
    }
    export_ = on;
    return true;
}

std::size_t Encoder::buffer_bytes(Buffer which) const {
    const std::size_t nblk{geo_.nblocks};
    const std::size_t tiles{static_cast<std::size_t>(geo_.cmw) * geo_.cmh};
    switch (which) {
        case Buffer::kCoeffs: return 3 * nblk * 64 * 2;
        case Buffer::kAcs: return nblk;
        case Buffer::kQuantField: return nblk * 4;
        case Buffer::kCflPartials: return static_cast<std::size_t>(geo_.rw) * geo_.rh * 12;
        case Buffer::kYtoX:
        case Buffer::kYtoB: return tiles;
        case Buffer::kAc: return 3 * nblk * 64 * 2;
        case Buffer::kDc: return 3 * nblk * 4;
        case Buffer::kNzGrid: return 3 * nblk;
        case Buffer::kNzCount: return 3 * nblk * 2;
        case Buffer::kHistograms: return (kNumClusters + 2 * geo_.num_dc) * kHistStride * 4;
        case Buffer::kAcGroupBytes: return geo_.num_ac * 4;
        case Buffer::kDcGroupBytes: return geo_.num_dc * 4;
        case Buffer::kCodeDepth: return (kNumClusters + 2 * geo_.num_dc) * kHistStride;
        case Buffer::kCodeBits: return (kNumClusters + 2 * geo_.num_dc) * kHistStride * 2;
        case Buffer::kFeatures: return features_ != nullptr ? static_cast<std::size_t>(geo_.rw) * geo_.rh * 108 * 4 : 0;
        default: return 0;
    }
}

bool Encoder::fetch(Buffer which, void* host, std::size_t bytes) {
    if (!configured_ || bytes > buffer_bytes(which) || hipSetDevice(device_) != hipSuccess) {
        return false;
    }
    const void* src{nullptr};
    switch (which) {
        case Buffer::kCoeffs: src = coeffs_; break;
        case Buffer::kAcs: src = acs_; break;
        case Buffer::kQuantField: src = qf_; break;
        case Buffer::kCflPartials: src = cfl_part_; break;
        case Buffer::kYtoX: src = ytox_; break;
        case Buffer::kYtoB: src = ytob_; break;
        case Buffer::kAc: src = ac_; break;
        case Buffer::kDc: src = dc_; break;
        case Buffer::kNzGrid: src = nz_grid_; break;
        case Buffer::kNzCount: src = nz_count_; break;
        case Buffer::kHistograms: src = hists_; break;
        case Buffer::kAcGroupBytes: src = ac_group_bytes_; break;
        case Buffer::kDcGroupBytes: src = dc_group_bytes_; break;
        case Buffer::kFeatures: src = features_; break;
        case Buffer::kCodeDepth:
        case Buffer::kCodeBits: {
            // Rows of the code set, one per histogram.
            const std::size_t nh{static_cast<std::size_t>(kNumClusters) + 2 * geo_.num_dc};
            const auto* cs{static_cast<const std::uint8_t*>(codes_)};
            const std::size_t row{static_cast<std::size_t>(which == Buffer::kCodeDepth ? kHistStride : kHistStride * 2)};
            const std::size_t base{static_cast<std::size_t>(which == Buffer::kCodeDepth ? 0 : 96 * kHistStride)};
            return hipMemcpy(host, cs + base, std::min(bytes, nh * row), hipMemcpyDeviceToHost) == hipSuccess;
        }
        default: return false;
    }
    return hipMemcpy(host, src, bytes, hipMemcpyDeviceToHost) == hipSuccess;
}

}  // namespace cjx
