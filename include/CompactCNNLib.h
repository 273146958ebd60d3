/*
 * CompactCNNLib.h -- C++ facade with the reference's public interface
 * (Bkmz21/CompactCNNCascade prebuild/include/CompactCNNLib.h:36-194), implemented
 * in libccnn.so over the C ABI of ccnn.h. Source-compatible with the reference's
 * examples/main.cpp. Differences: the precision enumerator spelled `default` in the
 * reference header (a C++ keyword) is `def`, as in the reference's own source header
 * (src/CNNObjectDetector/CompactCNNLibAPI_v2.h:52); no __declspec.
 *
 * As in the reference, every FaceDetector object shares one process-wide detector
 * guarded by one mutex (CompactCNNLibAPI_v2.cpp:32-33): a second Init or any
 * destructor clears the shared detector.
 */
#pragma once

#include <cstddef>

namespace CompactCNNLib
{
    class FaceDetector
    {
    public:
        enum struct Pipeline { CPU = 0, GPU = 1, GPU_CPU = 2 };
        enum struct DetectMode { disable = 0, sync = 1, async = 2 };
        enum struct DetectPrecision { def = 0, low = 1, normal = 2, high = 3, ultra = 4 };

        struct Size
        {
            int width = 0;
            int height = 0;
            Size() = default;
            Size(int _width, int _height) : width(_width), height(_height) { }
        };
        struct Face
        {
            int x = 0;
            int y = 0;
            int width = 0;
            int height = 0;
            float score = 0.f;
            int gender = -1;
            int smile = -1;
            int glasses = -1;
            Face() = default;
            Face(int _x, int _y, int _width, int _height, float _score = 0.f)
                : x(_x), y(_y), width(_width), height(_height), score(_score) { }
        };
        struct ImageData
        {
            int cols = 0;
            int rows = 0;
            int channels = 0;
            unsigned char* data = nullptr;   // 8-bit gray, BGR or BGRA
            size_t step = 0;
            ImageData() = default;
            ImageData(int _cols, int _rows, int _channels, unsigned char* _data, size_t _step)
                : cols(_cols), rows(_rows), channels(_channels), data(_data), step(_step) { }
        };
        struct Param
        {
            Size max_image_size = Size(1920, 1080);
            int min_face_height = 80;
            int max_face_height = 0;
            float scale_factor = 1.15f;
            int min_neighbors = 2;
            DetectPrecision detect_precision = DetectPrecision::normal;
            float treshold[3];
            bool drop_detect = true;
            bool equalize = true;
            bool reflection = true;
            Pipeline pipeline = Pipeline::GPU;
            DetectMode detect_mode = DetectMode::async;
            int num_threads = 2;
            const char* models[3];
            int index_output[3];
            bool device_info = true;
            int cuda_device_id = 0;          // HIP device ordinal
            int cl_platform_id = -1;
            int cl_device_id = -1;
            bool facial_analysis = false;
            Param()
            {
                treshold[0] = -1.5f;
                treshold[1] = -1.25f;
                treshold[2] = -1.0f;
                models[0] = nullptr;
                models[1] = nullptr;
                models[2] = nullptr;
                index_output[0] = 1;
                index_output[1] = 0;
                index_output[2] = 0;
            }
        };

        FaceDetector();
        ~FaceDetector();

        int Init(Param& param);
        int Detect(Face* faces, ImageData& img);
        int Clear();

        bool isEmpty() const;

        Param getParam() const;
        void setParam(Param& param);

        Size getMaxImageSize() const;
        void setMaxImageSize(Size max_image_size);

        int getMinObjectHeight() const;
        void setMinObjectHeight(int min_obj_height);

        int getMaxObjectHeight() const;
        void setMaxObjectHeight(int max_obj_height);

        int getNumThreads() const;
        void setNumThreads(int num_threads);

        int getGrayImage32F(ImageData* img, Pipeline pipeline);

        static void CNTKDump2Binary(const char* binary_file, const char* cntk_model_dump);
    };
}
