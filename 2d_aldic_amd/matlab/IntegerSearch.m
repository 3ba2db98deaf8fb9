function [DICpara,x0,y0,u,v,cc]= IntegerSearch(ImgRef,ImgDef,file_name,DICpara)
% Library-backed replacement of IntegerSearch.m:1-167 (called main_ALDIC.m:90): the
% whole-field FFT-free integer search of funIntegerSearch.m:82-200 (normxcorr2 cropped
% to the (2s+3)^2 lags, findpeak.m fit) on the GPU.  Non-interactive: the search radius
% is DICpara.SizeOfFFTSearchRegion (scalar or [sx sy]), the grid DICpara.gridxyROIRange,
% winsize and winstepsize as in DICpara.  Then, as IntegerSearch.m:68-76 and :114 do,
% nodes on masked pixels (DICpara.ImgRefMask) become NaN and the median test of
% funRemoveOutliers(u,v,cc,0.5,100) replaces the outliers (on the GPU, no prompts); the
% interactive bound/click clean-up of the reference is not part of this call.  With
% DICpara.InitFFTSearchMethod 0 the multigrid search of IntegerSearch.m:116-133 runs
% instead (funIntegerSearchMg.m, no outlier test, SizeOfFFTSearchRegion [0 0] after it).
% ImgRef / ImgDef are fNormalized / gNormalized.
grid = [DICpara.gridxyROIRange.gridx(1) DICpara.gridxyROIRange.gridx(2) ...
        DICpara.gridxyROIRange.gridy(1) DICpara.gridxyROIRange.gridy(2)];
aldic_mex('set_normalized', double(ImgRef), double(ImgDef));
[nx, ny] = size(ImgRef);
if isfield(DICpara, 'InitFFTSearchMethod') && isequal(DICpara.InitFFTSearchMethod, 0)
    [x0, y0, u, v, ccmax] = aldic_mex('integer_search_mg', grid, DICpara.winsize, DICpara.winstepsize);
    cc = struct('max', ccmax, 'qfactors', [], 'A', []);
    if isfield(DICpara, 'ImgRefMask') && ~isempty(DICpara.ImgRefMask)
        hole = ~DICpara.ImgRefMask(sub2ind([nx ny], x0(:), y0(:)));
        u(hole) = NaN; v(hole) = NaN;
    end
    DICpara.SizeOfFFTSearchRegion = [0 0];
    if nargin >= 3 && ~isempty(file_name), DICpara.ImgSize = [nx ny]; end
    return
end
s = DICpara.SizeOfFFTSearchRegion; if numel(s) == 1, s = [s s]; end
[x0, y0, u, v, ccmax, qf, ~] = aldic_mex('integer_search', nx, ny, grid, ...
    DICpara.winsize, DICpara.winstepsize, s, 1);
cc = struct('max', ccmax, 'qfactors', qf, 'A', []);
if isfield(DICpara, 'ImgRefMask') && ~isempty(DICpara.ImgRefMask) && ...
        all(x0(:) == round(x0(:))) && all(y0(:) == round(y0(:)))
    hole = ~DICpara.ImgRefMask(sub2ind([nx ny], x0(:), y0(:)));
    u(hole) = NaN; v(hole) = NaN;
end
[u, v] = aldic_mex('remove_outliers', u, v, 100);
DICpara.SizeOfFFTSearchRegion = s;
if nargin >= 3 && ~isempty(file_name), DICpara.ImgSize = [nx ny]; end
end
