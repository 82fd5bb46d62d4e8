// Headless counterpart of the reference's frame loop (source/VulkanPBRT.cpp:
// 314-618): builds the render modules through the drop-in API, records the
// trace -> accumulate -> denoise -> TAA -> copy-back command list once, then
// runs it per frame with updated push constants and camera matrices.
//
//   pbrt_frame [--config 1|2] [--seed S] [--width W --height H] [--frames F]
//              [--depth D] [--denoiser bmfr|bfr|learned|none] [--block 8|16|32|blend] [--taa]
//              [--graph] [--out final.ppm] [--raw image.bin]
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/pbrt_scenegen.h"
#include "../renderModules/RenderModules.hpp"

using namespace vkpbrt;

int main(int argc, char** argv) {
    uint32_t config = 1, frames = 4, depth = 2;
    uint64_t seed = 1;
    int width = 0, height = 0;
    std::string denoiser = "bmfr", block = "32", out, raw;
    bool taa = false, graph = false;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() { return i + 1 < argc ? std::string(argv[++i]) : std::string(); };
        if (a == "--config") config = std::stoul(next());
        else if (a == "--seed") seed = std::stoull(next());
        else if (a == "--width") width = std::stoi(next());
        else if (a == "--height") height = std::stoi(next());
        else if (a == "--frames") frames = std::stoul(next());
        else if (a == "--depth") depth = std::stoul(next());
        else if (a == "--denoiser") denoiser = next();
        else if (a == "--block") block = next();
        else if (a == "--taa") taa = true;
        else if (a == "--graph") graph = true;
        else if (a == "--out") out = next();
        else if (a == "--raw") raw = next();
        else {
            std::fprintf(stderr, "unknown argument %s\n", a.c_str());
            return 2;
        }
    }
    if (!width) width = config == 2 ? 1920 : 256;
    if (!height) height = config == 2 ? 1080 : 256;
    try {
        pbrt_generated_scene* gen = nullptr;
        check(pbrt_generate_scene(config, seed, &gen));
        pbrt_scene_desc desc;
        check(pbrt_generated_scene_desc(gen, &desc));
        pbrt_scene* scene = nullptr;
        check(pbrt_scene_create(&desc, &scene));
        hipStream_t stream;
        if (hipStreamCreate(&stream) != hipSuccess) throw Exception(PBRT_ERR_HIP, "hipStreamCreate");
        Context ctx{stream};

        auto push_constants = std::make_shared<PushConstants<RayTracingPushConstants>>();
        GBufferPtr g_buffer;
        IlluminationBufferPtr illumination_buffer;
        bool denoise = denoiser == "bmfr" || denoiser == "bfr" || denoiser == "learned";
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
        auto pbrt_pipeline = PBRTPipeline::create(scene, g_buffer, illumination_buffer, denoise,
                                                  RayTracingRayOrigin::CAMERA);
        pbrt_pipeline->max_recursion_depth = depth;
        pbrt_pipeline->set_tlas(scene);
        pbrt_pipeline->compile(ctx);
        auto commands = std::make_shared<Commands>();
        pbrt_pipeline->add_trace_rays_to_command_graph(commands, push_constants);

        std::shared_ptr<Accumulator> accumulator;
        DeviceImagePtr final_image;
        std::shared_ptr<Taa> taa_module;
        std::shared_ptr<FormatConverter> converter;
        if (denoise) {
            accumulator = Accumulator::create(g_buffer, pbrt_pipeline->get_illumination_buffer(), true);
            accumulator->compile_images(ctx);
            accumulator->add_dispatch_to_command_graph(commands);
            add_denoiser_to_commands(dtype, dsize, commands, ctx, width, height,
                                     push_constants, g_buffer, accumulator->accumulated_illumination,
                                     accumulator->accumulation_buffer, final_image);
            if (taa) {
                taa_module = Taa::create(width, height, 16, 16, g_buffer, accumulator->accumulation_buffer, final_image);
                taa_module->compile(ctx);
                taa_module->add_dispatch_to_command_graph(commands, push_constants);
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

        std::array<float, 16> prev_view{};
        std::vector<double> ms;
        hipEvent_t e0, e1;
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        for (uint32_t f = 0; f < frames; ++f) {
            pbrt_camera_pose pose;
            check(pbrt_generated_scene_camera(gen, f, &pose));
            double eye[3] = {pose.eye[0], pose.eye[1], pose.eye[2]};
            double center[3] = {pose.center[0], pose.center[1], pose.center[2]};
            double up[3] = {pose.up[0], pose.up[1], pose.up[2]};
            auto view = look_at(eye, center, up);
            auto proj = perspective(pose.fovy_degrees * 3.14159265358979323846 / 180.0, (double)width / height,
                                    pose.near_distance, pose.far_distance);
            auto& pc = push_constants->value();
            pc.view_inverse = to_float(inverse(view));
            pc.proj_inverse = to_float(inverse(proj));
            pc.prev_view = f == 0 ? to_float(view) : prev_view;
            pc.frame_number = f;
            pc.sample_number = denoise ? 0 : f;
            if (accumulator) {
                CameraMatrices cur, prev;
                cur.inv_view = pc.view_inverse;
                cur.inv_proj = pc.proj_inverse;
                cur.proj = to_float(proj);
                prev.view = pc.prev_view;
                accumulator->set_camera_matrices((int)f, cur, prev);
            }
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
            prev_view = to_float(view);
        }
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
        if (!raw.empty()) {
            std::ofstream f(raw, std::ios::binary);
            f.write(reinterpret_cast<const char*>(img.data()), (std::streamsize)img.size());
        }
        std::printf("{\"frames\": %u, \"width\": %d, \"height\": %d, \"ms\": [", frames, width, height);
        for (size_t i = 0; i < ms.size(); ++i) std::printf("%s%.3f", i ? ", " : "", ms[i]);
        std::printf("]}\n");
        commands.reset();
        (void)hipStreamDestroy(stream);
        check(pbrt_scene_destroy(scene));
        check(pbrt_generated_scene_destroy(gen));
    } catch (const Exception& e) {
        std::fprintf(stderr, "%s (status %d)\n", e.what(), e.status);
        return 1;
    }
    return 0;
}
