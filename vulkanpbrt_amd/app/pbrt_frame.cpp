// Headless counterpart of the reference's frame loop (source/VulkanPBRT.cpp:
// 314-618): builds the render modules through the drop-in API, records the
// trace -> accumulate -> denoise -> TAA -> copy-back command list once, then
// runs it per frame with updated push constants and camera matrices.
//
//   pbrt_frame [--config 1|2] [--seed S] [-i scene.obj] [--width W --height H] [--frames F | -f F]
//              [--depth D] [--denoiser bmfr|bfr|learned|none] [--block 8|16|32|blend] [--taa]
//              [--graph] [--out final.ppm] [--raw image.bin] [--png final.png] [--pfm final.pfm]
//              [--exportDepth|--exportPosition|--exportNormal|--exportAlbedo|--exportMaterial|
//               --exportIllumination fmt_%d.pfm] [--exportMatrices m.json]
//              [--depths|--positions fmt] --normals fmt --albedos fmt --illuminations fmt --matrices m.json
//
// Export and import follow VulkanPBRT.cpp:100-120,224-290,457-480 with PFM files in
// place of OpenEXR (include/pbrt_io.h).  With --normals the frames come from files
// (use_external_buffers): no scene is traced; the G-buffer and illumination of each
// frame are uploaded and accumulated, denoised and anti-aliased.  The statistics
// of the reference's GUI (source/Gui.hpp:39-44) are printed to stderr at the end.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/pbrt_io.h"
#include "../../include/pbrt_scenegen.h"
#include "../renderModules/RenderModules.hpp"

using namespace vkpbrt;

namespace {

std::string fmt_frame(const std::string& fmt, uint32_t f) {
    char buf[512];
    std::snprintf(buf, sizeof(buf), fmt.c_str(), (int)f);
    return buf;
}

template <class T>
std::vector<T> download(const DeviceImagePtr& img) {
    std::vector<T> v(img->bytes() / sizeof(T));
    if (hipMemcpy(v.data(), img->data(), img->bytes(), hipMemcpyDeviceToHost) != hipSuccess)
        throw Exception(PBRT_ERR_HIP, "hipMemcpy to host");
    return v;
}

template <class T>
void upload(const DeviceImagePtr& img, const std::vector<T>& v) {
    if (v.size() * sizeof(T) != img->bytes()) throw Exception(PBRT_ERR_INVALID_ARG, "image size does not match");
    if (hipMemcpy(img->data(), v.data(), img->bytes(), hipMemcpyHostToDevice) != hipSuccess)
        throw Exception(PBRT_ERR_HIP, "hipMemcpy to device");
}

std::vector<float> read_pfm(const std::string& path, uint32_t w, uint32_t h, uint32_t want_channels) {
    uint32_t fw = 0, fh = 0, ch = 0;
    float* px = nullptr;
    check(pbrt_pfm_read(path.c_str(), &fw, &fh, &ch, &px));
    std::vector<float> v(px, px + (size_t)fw * fh * ch);
    pbrt_io_free(px);
    if (fw != w || fh != h || ch != want_channels)
        throw Exception(PBRT_ERR_INVALID_ARG, path + ": unexpected size or channel count");
    return v;
}

// rgb -> rgba with a constant alpha
std::vector<float> with_alpha(const std::vector<float>& rgb, float a) {
    std::vector<float> v(rgb.size() / 3 * 4);
    for (size_t i = 0; i < rgb.size() / 3; ++i) {
        v[4 * i] = rgb[3 * i];
        v[4 * i + 1] = rgb[3 * i + 1];
        v[4 * i + 2] = rgb[3 * i + 2];
        v[4 * i + 3] = a;
    }
    return v;
}

pbrt_camera_matrices to_c(const CameraMatrices& m) {
    pbrt_camera_matrices c{};
    std::memcpy(c.view, m.view.data(), 64);
    std::memcpy(c.inv_view, m.inv_view.data(), 64);
    if (m.proj && m.inv_proj) {
        std::memcpy(c.proj, m.proj->data(), 64);
        std::memcpy(c.inv_proj, m.inv_proj->data(), 64);
        c.has_proj = 1;
    }
    return c;
}

CameraMatrices from_c(const pbrt_camera_matrices& c) {
    CameraMatrices m;
    std::memcpy(m.view.data(), c.view, 64);
    std::memcpy(m.inv_view.data(), c.inv_view, 64);
    if (c.has_proj) {
        std::array<float, 16> p, ip;
        std::memcpy(p.data(), c.proj, 64);
        std::memcpy(ip.data(), c.inv_proj, 64);
        m.proj = p;
        m.inv_proj = ip;
    }
    return m;
}

}  // namespace

int main(int argc, char** argv) {
    uint32_t config = 1, frames = 4, depth = 2, spp = 1;
    uint64_t seed = 1;
    int width = 0, height = 0;
    std::string denoiser = "bmfr", block = "32", out, raw, png, pfm, scene_file;
    // VulkanPBRT.cpp:100-113 file arguments
    std::string depths, positions, normals, albedos, materials, illuminations, matrices_path;
    std::string ex_depth, ex_position, ex_normal, ex_albedo, ex_material, ex_illumination, ex_matrices;
    bool taa = false, graph = false;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() { return i + 1 < argc ? std::string(argv[++i]) : std::string(); };
        if (a == "--config") config = std::stoul(next());
        else if (a == "--seed") seed = std::stoull(next());
        else if (a == "-i") scene_file = next();
        else if (a == "--width") width = std::stoi(next());
        else if (a == "--height") height = std::stoi(next());
        else if (a == "--frames" || a == "-f") frames = std::stoul(next());
        else if (a == "--spp") spp = std::stoul(next());
        else if (a == "--depth") depth = std::stoul(next());
        else if (a == "--denoiser") denoiser = next();
        else if (a == "--block") block = next();
        else if (a == "--taa") taa = true;
        else if (a == "--graph") graph = true;
        else if (a == "--out") out = next();
        else if (a == "--raw") raw = next();
        else if (a == "--png") png = next();
        else if (a == "--pfm") pfm = next();
        else if (a == "--depths") depths = next();
        else if (a == "--positions") positions = next();
        else if (a == "--normals") normals = next();
        else if (a == "--albedos") albedos = next();
        else if (a == "--materials") materials = next();
        else if (a == "--illuminations") illuminations = next();
        else if (a == "--matrices") matrices_path = next();
        else if (a == "--exportDepth") ex_depth = next();
        else if (a == "--exportPosition") ex_position = next();
        else if (a == "--exportNormal") ex_normal = next();
        else if (a == "--exportAlbedo") ex_albedo = next();
        else if (a == "--exportMaterial") ex_material = next();
        else if (a == "--exportIllumination") ex_illumination = next();
        else if (a == "--exportMatrices") ex_matrices = next();
        else {
            std::fprintf(stderr, "unknown argument %s\n", a.c_str());
            return 2;
        }
    }
    if (!width) width = config == 2 ? 1920 : 256;
    if (!height) height = config == 2 ? 1080 : 256;
    const bool use_external_buffers = !normals.empty();
    const bool export_illumination = !ex_illumination.empty();
    const bool export_g_buffer = !ex_depth.empty() || !ex_position.empty() || !ex_normal.empty() ||
                                 !ex_albedo.empty() || !ex_material.empty();
    const bool store_matrices = export_g_buffer || !ex_matrices.empty();
    if (spp == 0) spp = 1;
    try {
        const uint32_t W = (uint32_t)width, H = (uint32_t)height, npx = W * H;
        std::vector<pbrt_camera_matrices> file_matrices;
        if (use_external_buffers) {
            if (matrices_path.empty() || illuminations.empty() || albedos.empty() ||
                (depths.empty() && positions.empty()))
                throw Exception(PBRT_ERR_INVALID_ARG,
                                "external buffers need --matrices, --illuminations, --albedos and --depths or "
                                "--positions with --normals");
            pbrt_camera_matrices* m = nullptr;
            uint32_t n = 0;
            check(pbrt_matrices_read(matrices_path.c_str(), &m, &n));
            file_matrices.assign(m, m + n);
            pbrt_io_free(m);
            if (file_matrices.size() < frames) throw Exception(PBRT_ERR_INVALID_ARG, "fewer matrices than frames");
        }
        pbrt_generated_scene* gen = nullptr;
        pbrt_loaded_scene* loaded = nullptr;
        pbrt_scene* scene = nullptr;
        uint32_t triangles = 0;
        if (!use_external_buffers) {
            check(pbrt_generate_scene(config, seed, &gen));  // the camera path (and the scene without -i)
            pbrt_scene_desc desc;
            check(pbrt_generated_scene_desc(gen, &desc));
            if (!scene_file.empty()) {
                const pbrt_light* lights = desc.lights;
                const uint32_t nl = desc.light_count;
                check(pbrt_load_obj(scene_file.c_str(), nullptr, &loaded));
                check(pbrt_loaded_scene_desc(loaded, &desc));
                desc.lights = lights;  // the workload's explicit lights; the file carries none
                desc.light_count = nl;
            }
            check(pbrt_scene_create(&desc, &scene));
            pbrt_scene_info info;
            check(pbrt_scene_get_info(scene, &info));
            triangles = info.triangle_count;
        }
        hipStream_t stream;
        if (hipStreamCreate(&stream) != hipSuccess) throw Exception(PBRT_ERR_HIP, "hipStreamCreate");
        Context ctx{stream};

        auto push_constants = std::make_shared<PushConstants<RayTracingPushConstants>>();
        GBufferPtr g_buffer;
        IlluminationBufferPtr illumination_buffer;
        bool denoise = denoiser == "bmfr" || denoiser == "bfr" || denoiser == "learned";
        if (use_external_buffers && !denoise)
            throw Exception(PBRT_ERR_INVALID_ARG, "external buffers are denoised: pick a denoiser");
        const DenoisingType dtype = denoiser == "bfr"       ? DenoisingType::BFR
                                    : denoiser == "learned" ? DenoisingType::LEARNED
                                                            : DenoisingType::BMFR;
        const DenoisingBlockSize dsize = block == "8"    ? DenoisingBlockSize::X8
                                         : block == "16" ? DenoisingBlockSize::X16
                                         : block == "blend" ? DenoisingBlockSize::X8X16X32
                                                            : DenoisingBlockSize::X32;
        if (denoise) {
            g_buffer = GBuffer::create(width, height);
            illumination_buffer = IlluminationBufferDemodulatedFloat(width, height);
        } else {
            illumination_buffer = IlluminationBufferFinalFloat(width, height);
        }
        auto commands = std::make_shared<Commands>();
        std::shared_ptr<PBRTPipeline> pbrt_pipeline;
        if (!use_external_buffers) {
            pbrt_pipeline = PBRTPipeline::create(scene, g_buffer, illumination_buffer, denoise,
                                                 RayTracingRayOrigin::CAMERA);
            pbrt_pipeline->max_recursion_depth = depth;
            pbrt_pipeline->set_tlas(scene);
            pbrt_pipeline->compile(ctx);
            pbrt_pipeline->add_trace_rays_to_command_graph(commands, push_constants);
            illumination_buffer = pbrt_pipeline->get_illumination_buffer();
        } else {
            g_buffer->compile(ctx);
            illumination_buffer->compile(ctx);
        }

        std::shared_ptr<Accumulator> accumulator;
        DeviceImagePtr final_image;
        std::shared_ptr<Taa> taa_module;
        std::shared_ptr<FormatConverter> converter;
        if (denoise) {
            accumulator = Accumulator::create(g_buffer, illumination_buffer, true);
            accumulator->compile_images(ctx);
            accumulator->add_dispatch_to_command_graph(commands);
            add_denoiser_to_commands(dtype, dsize, commands, ctx, width, height,
                                     push_constants, g_buffer, accumulator->accumulated_illumination,
                                     accumulator->accumulation_buffer, final_image);
            if (taa) {
                taa_module = Taa::create(width, height, 16, 16, g_buffer, accumulator->accumulation_buffer, final_image);
                taa_module->compile(ctx);
                taa_module->add_dispatch_to_command_graph(commands);
                final_image = taa_module->get_final_descriptor_image();
            }
            accumulator->accumulation_buffer->copy_to_back_images(commands, g_buffer,
                                                                  accumulator->accumulated_illumination);
        } else {
            converter = std::make_shared<FormatConverter>(illumination_buffer->illumination_images[0], Format::BGRA8);
            converter->compile_images(ctx);
            converter->add_dispatch_to_command_graph(commands);
            final_image = converter->final_image;
        }
        if ((export_g_buffer || export_illumination) && !denoise)
            throw Exception(PBRT_ERR_INVALID_ARG, "G-buffer and illumination export need a denoiser's buffers");

        std::array<float, 16> prev_view{};
        std::vector<double> ms;
        std::vector<pbrt_camera_matrices> export_matrices;
        hipEvent_t e0, e1;
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        // sample_index restarts per frame index: VulkanPBRT.cpp:551-617
        uint32_t frame_index = 0, sample_index = 0;
        while (frame_index < frames) {
            const uint32_t f = frame_index;
            auto& pc = push_constants->value();
            CameraMatrices cur, prev;
            if (use_external_buffers) {
                cur = from_c(file_matrices[f]);
                prev = from_c(file_matrices[f != 0 ? f - 1 : f]);
                pc.view_inverse = cur.inv_view;
                if (cur.inv_proj) pc.proj_inverse = *cur.inv_proj;
                pc.prev_view = prev.view;
                // OfflineGBuffer::upload_to_g_buffer_command and the illumination upload
                std::vector<float> d;
                if (!depths.empty()) {
                    d = read_pfm(fmt_frame(depths, f), W, H, 1);
                } else {
                    const std::vector<float> p = with_alpha(read_pfm(fmt_frame(positions, f), W, H, 3), 1.0f);
                    d.resize(npx);
                    check(pbrt_position_to_depth(p.data(), npx, &file_matrices[f], d.data()));
                }
                upload(g_buffer->depth, d);
                const std::vector<float> n = with_alpha(read_pfm(fmt_frame(normals, f), W, H, 3), 1.0f);
                std::vector<float> sph(2 * (size_t)npx);
                check(pbrt_normals_to_spherical(n.data(), npx, sph.data()));
                upload(g_buffer->normal, sph);
                const std::vector<float> al = with_alpha(read_pfm(fmt_frame(albedos, f), W, H, 3), 1.0f);
                std::vector<uint8_t> a8(4 * (size_t)npx);
                check(pbrt_albedo_from_float(al.data(), npx, a8.data()));
                upload(g_buffer->albedo, a8);
                upload(illumination_buffer->illumination_images[0],
                       with_alpha(read_pfm(fmt_frame(illuminations, f), W, H, 3), 1.0f));
            } else {
                pbrt_camera_pose pose;
                check(pbrt_generated_scene_camera(gen, f, &pose));
                double eye[3] = {pose.eye[0], pose.eye[1], pose.eye[2]};
                double center[3] = {pose.center[0], pose.center[1], pose.center[2]};
                double up[3] = {pose.up[0], pose.up[1], pose.up[2]};
                auto view = look_at(eye, center, up);
                auto proj = perspective(pose.fovy_degrees * 3.14159265358979323846 / 180.0, (double)width / height,
                                        pose.near_distance, pose.far_distance);
                pc.view_inverse = to_float(inverse(view));
                pc.proj_inverse = to_float(inverse(proj));
                pc.prev_view = f == 0 && sample_index == 0 ? to_float(view) : prev_view;
                cur.view = to_float(view);
                cur.inv_view = pc.view_inverse;
                cur.inv_proj = pc.proj_inverse;
                cur.proj = to_float(proj);
                prev.view = pc.prev_view;
                prev_view = to_float(view);
            }
            pc.frame_number = f;
            pc.sample_number = denoise ? 0 : f * spp + sample_index;
            if (accumulator) accumulator->set_camera_matrices((int)f, cur, prev);
            (void)hipEventRecord(e0, stream);
            if (graph)
                commands->launch_graph(stream);
            else
                commands->submit(stream);
            (void)hipEventRecord(e1, stream);
            (void)hipStreamSynchronize(stream);
            float t = 0;
            (void)hipEventElapsedTime(&t, e0, e1);
            ms.push_back(t);
            if (sample_index + 1 >= spp) {
                // GBufferIO::export_g_buffer / IlluminationBufferIO::export_illumination per frame
                // (the G-buffer after the frame: copy_to_back_images keeps it in place)
                if (export_g_buffer && !use_external_buffers) {
                    const auto mats = to_c(cur);
                    if (!ex_depth.empty() || !ex_position.empty()) {
                        const auto d = download<float>(g_buffer->depth);
                        if (!ex_depth.empty()) check(pbrt_pfm_write(fmt_frame(ex_depth, f).c_str(), W, H, 1, d.data(), 1));
                        if (!ex_position.empty()) {
                            std::vector<float> p(4 * (size_t)npx);
                            check(pbrt_depth_to_position(d.data(), W, H, &mats, p.data()));
                            check(pbrt_pfm_write(fmt_frame(ex_position, f).c_str(), W, H, 3, p.data(), 4));
                        }
                    }
                    if (!ex_normal.empty()) {
                        const auto sph = download<float>(g_buffer->normal);
                        std::vector<float> n(4 * (size_t)npx);
                        check(pbrt_spherical_to_normals(sph.data(), npx, n.data()));
                        check(pbrt_pfm_write(fmt_frame(ex_normal, f).c_str(), W, H, 3, n.data(), 4));
                    }
                    for (auto [img, path] : {std::make_pair(g_buffer->albedo, ex_albedo),
                                             std::make_pair(g_buffer->material, ex_material)}) {
                        if (path.empty()) continue;
                        const auto u8 = download<uint8_t>(img);
                        std::vector<float> v(4 * (size_t)npx);
                        check(pbrt_unorm_to_float(u8.data(), npx, v.data()));
                        check(pbrt_pfm_write(fmt_frame(path, f).c_str(), W, H, 3, v.data(), 4));
                    }
                }
                if (export_illumination && !use_external_buffers) {
                    const auto il = download<float>(illumination_buffer->illumination_images[0]);
                    check(pbrt_pfm_write(fmt_frame(ex_illumination, f).c_str(), W, H, 3, il.data(), 4));
                }
                if (store_matrices) export_matrices.push_back(to_c(cur));
                ++frame_index;
                sample_index = 0;
            } else {
                ++sample_index;
            }
        }
        if (!ex_matrices.empty())
            check(pbrt_matrices_write(ex_matrices.c_str(), export_matrices.data(), (uint32_t)export_matrices.size()));
        std::vector<uint8_t> img((size_t)width * height * 4);
        (void)hipMemcpy(img.data(), final_image->data(), img.size(), hipMemcpyDeviceToHost);
        if (!out.empty()) {
            std::ofstream f(out, std::ios::binary);
            f << "P6\n" << width << " " << height << "\n255\n";
            for (size_t i = 0; i < (size_t)width * height; ++i) {
                const uint8_t* p = &img[4 * i];  // B8G8R8A8
                f.put((char)p[2]).put((char)p[1]).put((char)p[0]);
            }
        }
        if (!png.empty()) check(pbrt_image_write_png(png.c_str(), W, H, img.data(), 1));
        if (!pfm.empty()) {
            std::vector<float> rgb(3 * (size_t)npx);
            for (size_t i = 0; i < npx; ++i)
                for (int c = 0; c < 3; ++c) rgb[3 * i + c] = (float)img[4 * i + 2 - c] / 255.0f;
            check(pbrt_pfm_write(pfm.c_str(), W, H, 3, rgb.data(), 3));
        }
        if (!raw.empty()) {
            std::ofstream f(raw, std::ios::binary);
            f.write(reinterpret_cast<const char*>(img.data()), (std::streamsize)img.size());
        }
        // the reference GUI's statistics (Gui.hpp:39-44): rays per pixel = 2 * max depth
        double sum = 0;
        for (double t : ms) sum += t;
        const double avg = ms.empty() ? 0.0 : sum / ms.size(), fps = avg > 0 ? 1000.0 / avg : 0.0;
        const int rays_per_pixel = (int)depth * 2;
        std::fprintf(stderr, "Application average %.3f ms/frame (%.1f FPS) for %d triangles\n", avg, fps,
                     (int)triangles);
        std::fprintf(stderr, "Render size is %d by %d with %d rays/pixel resulting in %.3f mRays/second\n", width,
                     height, rays_per_pixel, fps * rays_per_pixel * width * height / 1.0e6);
        std::fprintf(stderr, "Samples per pixel: %u\n", spp);
        std::printf("{\"frames\": %u, \"width\": %d, \"height\": %d, \"ms\": [", frames, width, height);
        for (size_t i = 0; i < ms.size(); ++i) std::printf("%s%.3f", i ? ", " : "", ms[i]);
        std::printf("]}\n");
        commands.reset();
        pbrt_pipeline.reset();
        (void)hipStreamDestroy(stream);
        if (scene) check(pbrt_scene_destroy(scene));
        if (loaded) check(pbrt_loaded_scene_destroy(loaded));
        if (gen) check(pbrt_generated_scene_destroy(gen));
    } catch (const Exception& e) {
        std::fprintf(stderr, "%s (status %d)\n", e.what(), e.status);
        return 1;
    }
    return 0;
}
